"""
NaN handling on the GPU: the scenarios and expected answers of the
reference's find_nan tests (``myokit/tests/test_simulation_opencl.py``,
``SimulationOpenCLFindNanTest``, lines 1448-1589), restated.

The voltage-clamped model: LR1991 with V demoted to a bound variable (the
pacing level), Ca_i labelled as the membrane potential, V moved to the
engine component. At V = -47.13 mV, ina.m's alpha is 0/0, so the cell
clamped there by the protocol turns NaN at the first step after 1.234 ms.

These simulations run in double precision. In single precision the kernel
writes exp(k * (V + c)) as exp2(fma(V, k log2(e), c k log2(e))), which does
not give exactly exp(0) = 1 at V = -c: the removable singularity yields a
finite value there instead of the reference's 0/0 = NaN.
"""
import numpy as np
import pytest

import myokit_amd as myokit

pytestmark = pytest.mark.gpu


def clamped_model():
    m = myokit.load_model('example')
    m.binding('pace').set_binding(None)
    v = m.get('membrane.V')
    v.set_rhs(-80)
    v.demote()
    v.set_label(None)
    v.set_binding('pace')
    m.get('ica.Ca_i').set_label('membrane_potential')
    m.get('membrane').move_variable(v, m.get('engine'))
    return m


def protocol(t=1.234):
    p = myokit.Protocol()
    p.schedule(start=0, level=-80, duration=t)
    p.schedule(start=t, level=-47.13, duration=1000)
    return p


@pytest.fixture(scope='module')
def s2():
    s = myokit.SimulationOpenCL(clamped_model(), protocol(), ncells=(3, 3),
                                precision=myokit.DOUBLE_PRECISION)
    s.set_paced_cell_list([[1, 1]])
    return s


def test_big_steps():
    s = myokit.SimulationOpenCL(myokit.load_model('example'),
                                myokit.pacing.blocktrain(1000, 2), ncells=10)
    s.set_step_size(1)
    with pytest.raises(myokit.SimulationError):
        s.run(10)
    for log in (['membrane.V'], myokit.LOG_STATE, myokit.LOG_BOUND):
        s.reset()
        with pytest.raises(myokit.SimulationError, match='Unable to pinpoint'):
            s.run(10, log=log)


def test_one_over_zero(s2):
    s2.reset()
    with pytest.raises(myokit.SimulationError, match='Time: +1.23') as e:
        s2.run(5)
    print(str(e.value)[:600], flush=True)


def test_first_point(s2):
    s2.reset()
    x = s2.state(0, 0)
    x[3] = float('nan')
    s2.set_state(x, 2, 2)
    with pytest.raises(myokit.SimulationError, match='met in the very first'):
        s2.run(2)


def test_without_error(s2):
    try:
        s2.set_protocol(None)
        s2.reset()
        d = s2.run(10)
        with pytest.raises(myokit.FindNanError, match='not found in log'):
            s2.find_nan(d)
    finally:
        s2.set_protocol(protocol())


def test_manual_call(s2):
    p = myokit.Protocol()
    p.schedule(start=0, level=-90, duration=4)
    p.schedule(start=4, level=-47.13, duration=1000)
    try:
        s2.set_protocol(p)
        s2.reset()
        x = s2.state(0, 0)
        x[2] = 0.9      # lower j in one cell (makes the bisection branch)
        s2.set_state(x, 2, 2)
        d = s2.run(10, report_nan=False)
        time, icell, variable, value, states, bounds = s2.find_nan(d)
        assert time == pytest.approx(4.005, abs=1e-5)
        assert icell == [1, 1]
        assert variable == 'ina.m'
        with pytest.raises(myokit.FindNanError, match='not found'):
            s2.find_nan(d, 'x.y', [0, 1])
        with pytest.raises(myokit.FindNanError, match='state'):
            s2.find_nan(d, 'engine.time', [0, 1])
        with pytest.raises(myokit.FindNanError, match='safe range'):
            s2.find_nan(d, 'ina.m')
        with pytest.raises(myokit.FindNanError, match='lower than'):
            s2.find_nan(d, 'ina.m', [1, 0])
        # Unpaced cells sit at V = 0, so ina inactivates quickly
        time, icell, variable, value, states, bounds = s2.find_nan(
            d, 'ina.j', [0.5, 1])
        assert 2 < time < 3
        assert icell == [2, 2]
        assert variable == 'ina.j'
        time, icell, variable, value, states, bounds = s2.find_nan(
            d, 'ina.j', [0.2, 1])
        assert 5 < time < 6
        assert icell == [2, 2]
        assert variable == 'ina.j'
        assert len(states) == 4 and len(bounds) == 4
        assert np.isfinite(states[1]).all()
    finally:
        s2.set_protocol(protocol())


def test_nan_in_user_cell_0_of_a_reordered_graph():
    # The NaN test reads the caller's cell 0 (openclsim.c:1087-1089), also
    # when the device holds a connection graph in another cell order. Here
    # cell 0 and cell 1 form one component and cells 2..399 a random
    # geometric graph; only cell 0 is paced, so only cells 0 and 1 go NaN
    # (at the first step after 1.234 ms). The device cell at index 0 lies
    # in the other component: a test of device cell 0 would never halt.
    from oracle.cgen import OracleSimulation
    rng = np.random.default_rng(7)
    n = 400
    pts = rng.uniform(size=(n, 2))
    r2 = 6 / (np.pi * n)
    conns = [(0, 1, 5.0)]
    for i in range(2, n):
        d2 = np.sum((pts[i + 1:] - pts[i]) ** 2, axis=1)
        for j in np.nonzero(d2 < r2)[0]:
            conns.append((i, i + 1 + int(j), float(rng.uniform(5, 12))))
    s = myokit.SimulationOpenCL(clamped_model(), protocol(), ncells=n,
                                precision=myokit.DOUBLE_PRECISION)
    s.set_connections(conns)
    s.set_paced_cell_list([0])
    perm, iperm = s._cell_order()
    assert perm is not None and perm[0] > 1
    # State 0 of the clamped model (the one the NaN test reads) is ina.m;
    # the diffusing "membrane potential" is Ca_i
    qs = ['ina.m', 'ica.Ca_i']
    keys = ['engine.time'] + [str(c) + '.' + q for q in qs for c in range(n)]
    print('graph NaN: 400 cells,', len(conns), 'connections', flush=True)
    d = s.run(5, log=keys, log_interval=0.1, report_nan=False)
    o = OracleSimulation(clamped_model(), protocol(), ncells=n)
    o.connections = conns
    o.set_paced_cell_list([0])
    o.dt = s._step_size
    r = o.run(5, log_vars=qs, log_interval=0.1)
    assert r['halted']
    rows = len(d['engine.time'])
    assert rows == r['rows'] and rows < 20
    np.testing.assert_array_equal(np.asarray(d['engine.time']), r['time'])
    for q in qs:
        V = np.stack([np.asarray(d[str(c) + '.' + q]) for c in range(n)],
                     axis=1)
        np.testing.assert_array_equal(V, r[q])
        # only the paced cell 0 and its partner cell 1 can go NaN
        assert np.isfinite(V[:, 2:]).all()
        assert np.isfinite(V[:-1]).all()
    # the run halts at the first log row where the caller's cell 0 has a
    # NaN state 0, and at no earlier row
    m0 = np.asarray(d['0.ina.m'])
    assert np.isnan(m0[-1]) and np.isfinite(m0[:-1]).all()
    # find_nan needs every state and bound variable: the default log
    s.reset()
    d = s.run(5, log_interval=0.1, report_nan=False)
    assert len(d['engine.time']) == rows
    time, icell, variable, value, states, bounds = s.find_nan(d)
    assert time == pytest.approx(1.235, abs=1e-9)
    assert icell == [0]
    s.reset()
    with pytest.raises(myokit.SimulationError, match='Time: +1.23'):
        s.run(5, log_interval=0.1)
