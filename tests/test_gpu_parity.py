"""
GPU parity tests: SimulationOpenCL (HIP, MI355X) against the CPU oracle and
against the reference's Simulation1d golden runs.

Tolerances: double precision must match the oracle to max|dV| <= 1e-9 mV
(the engines use the same operation order and the same portable math, so
the expected difference is zero; the requirement is 1e-6 mV). Single
precision is compared with the double-precision oracle: activation times
within 0.1 ms and V RMS error <= 0.5 mV.
"""
import os

import numpy as np
import pytest

import myokit_amd as myokit
from helpers import Scenario, activation_times, run_both

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), 'golden')
FP64_TOL = 1e-9


def _report(name, P, O):
    d = np.max(np.abs(P['V'] - O['V'])) if P['V'].size else 0.0
    print(name, 'rows', P['V'].shape, 'max|dV| =', d, flush=True)
    return d


@pytest.mark.parametrize('case', ['fe', 'rl'])
def test_fp64_cable_lr1991(case):
    g = np.load(os.path.join(GOLD, 'sim1d_lr1991_' + case + '.npz'))
    n, npaced, gx, rl, dt, dur, li = g['config']
    sc = Scenario('lr1991', int(n), rl=bool(rl), dt=dt, gx=gx,
                  events=[tuple(g['event'])], paced=(int(npaced),))
    P, O = run_both(sc, dur, myokit.DOUBLE_PRECISION, log_interval=li)
    assert P['V'].shape == O['V'].shape == g['V'].shape
    assert _report('lr1991 ' + case, P, O) <= FP64_TOL
    np.testing.assert_array_equal(P['time'], O['time'])
    # Against the reference Simulation1d
    assert np.max(np.abs(P['V'] - g['V'])) < 1e-9
    assert np.max(np.abs(P['state'] - O['state'])) <= FP64_TOL


def test_fp64_config1_geometry():
    # 200-cell LR1991 cable, FE, dt 0.005, gx 9.5, 5 paced cells
    g = np.load(os.path.join(GOLD, 'sim1d_lr1991_cfg1.npz'))
    sc = Scenario('lr1991', 200, dt=0.005, gx=9.5, events=[(1, 0, 2, 1000, 0)],
                  paced=(5,))
    P, O = run_both(sc, 60.0, myokit.DOUBLE_PRECISION, log_interval=1.0)
    assert _report('config1 60ms', P, O) <= FP64_TOL
    assert np.max(np.abs(P['V'] - g['V'])) < 1e-9


def test_fp64_br1977():
    g = np.load(os.path.join(GOLD, 'sim1d_br1977_fe.npz'))
    sc = Scenario('br1977', 10, dt=0.005, gx=10.0, events=[(1, 1, 2, 1000, 0)],
                  paced=(2,))
    P, O = run_both(sc, 15.0, myokit.DOUBLE_PRECISION, log_interval=0.5)
    assert _report('br1977', P, O) <= FP64_TOL
    assert np.max(np.abs(P['V'] - g['V'])) < 1e-9


@pytest.mark.parametrize('shape', [(24, 16), (48, 40)])
def test_fp64_sheet_rl(shape):
    # Small sheet runs in one workgroup; the larger one on the grid driver
    sc = Scenario('lr1991', shape, rl=True, dt=0.01, gx=9.5, gy=9.5,
                  events=[(1, 0, 2, 1000, 0)], paced=(4, shape[1], 0, 0))
    P, O = run_both(sc, 15.0, myokit.DOUBLE_PRECISION, log_interval=0.5)
    assert _report('sheet ' + str(shape), P, O) <= FP64_TOL
    assert np.max(np.abs(P['state'] - O['state'])) <= FP64_TOL


def test_fp64_conductance_field():
    rng = np.random.default_rng(5)
    nx, ny = 20, 12
    gx = rng.uniform(5, 12, size=(ny, nx - 1))
    gy = rng.uniform(5, 12, size=(ny - 1, nx))
    sc = Scenario('lr1991', (nx, ny), rl=True, dt=0.01,
                  events=[(1, 0, 2, 1000, 0)], paced=(3, 3, 0, 0),
                  conductance_field=(gx, gy))
    P, O = run_both(sc, 12.0, myokit.DOUBLE_PRECISION)
    assert _report('hetero', P, O) <= FP64_TOL


def test_fp64_connections_graph():
    # Two chains and a random graph component; summation order follows the
    # connection list
    rng = np.random.default_rng(7)
    n = 120
    conns = [(i, i + 1, 9.5) for i in range(0, 39)]
    conns += [(i + 1, i, 7.0) for i in range(40, 79)]
    for k in range(60):
        i, j = rng.choice(np.arange(80, 120), size=2, replace=False)
        if (min(i, j), max(i, j)) not in {(min(a, b), max(a, b))
                                          for a, b, _ in conns}:
            conns.append((int(i), int(j), float(rng.uniform(5, 12))))
    sc = Scenario('lr1991', n, dt=0.005, events=[(1, 0, 2, 1000, 0)],
                  paced_list=[0, 1, 2, 40, 41, 80, 81], connections=conns)
    P, O = run_both(sc, 20.0, myokit.DOUBLE_PRECISION)
    assert _report('graph', P, O) <= FP64_TOL


def test_fp64_connections_large_graph_grid_driver():
    # A component larger than a workgroup's LDS: global CSR gather
    n = 3000
    conns = [(i, i + 1, 9.5) for i in range(n - 1)]
    sc = Scenario('lr1991', n, dt=0.005, events=[(1, 0, 2, 1000, 0)],
                  paced_list=[0, 1, 2], connections=conns)
    P, O = run_both(sc, 8.0, myokit.DOUBLE_PRECISION, log_interval=1.0)
    assert _report('graph-grid', P, O) <= FP64_TOL


def test_connections_equal_stencil():
    # A chain given as connections equals the homogeneous 1D stencil
    # (test_simulation_opencl.py:460-528 requires <= 1e-9 in double)
    n = 30
    a = Scenario('lr1991', n, dt=0.005, gx=9.5, paced=(3,))
    b = Scenario('lr1991', n, dt=0.005, paced_list=[0, 1, 2],
                 connections=[(i, i + 1, 9.5) for i in range(n - 1)])
    sa, sb = a.product(64), b.product(64)
    da = sa.run(10, log=['membrane.V'], log_interval=0.5)
    db = sb.run(10, log=['membrane.V'], log_interval=0.5)
    for k in da.keys():
        assert np.max(np.abs(np.asarray(da[k]) - np.asarray(db[k]))) <= 1e-9


def test_fp64_uncoupled_fields():
    rng = np.random.default_rng(11)
    n = 300
    s = np.exp(rng.uniform(np.log(0.5), np.log(2), size=(3, n)))
    fields = {'ina.gNa': 16 * s[0], 'ica.gCa': 0.09 * s[1],
              'ik.IK.gK': 0.282 * s[2]}
    sc = Scenario('lr1991', n, diffusion=False, dt=0.005,
                  events=[(1, 0, 2, 1000, 0)], fields=fields)
    P, O = run_both(sc, 40.0, myokit.DOUBLE_PRECISION, log_interval=1.0)
    assert _report('uncoupled fields', P, O) <= FP64_TOL
    # Cells differ through their fields
    assert np.std(P['V'][-1]) > 0.01


def test_uncoupled_identical_without_fields():
    # diffusion=False and no fields: every cell identical, bit for bit
    # (test_simulation_opencl.py:602-658)
    s = myokit.SimulationOpenCL(myokit.load_model('example'),
                                myokit.pacing.blocktrain(1000, 2), ncells=100,
                                diffusion=False, precision=32)
    d = s.run(20, log=['membrane.V'], log_interval=1)
    V = np.stack([d[str(i) + '.membrane.V'] for i in range(100)], axis=1)
    assert np.all(V == V[:, :1])


@pytest.mark.parametrize('rl', [False, True])
def test_fp32_against_fp64_oracle(rl):
    dt = 0.01 if rl else 0.005
    sc = Scenario('lr1991', 60, rl=rl, dt=dt, gx=9.5,
                  events=[(1, 0, 2, 1000, 0)], paced=(5,))
    P, O = run_both(sc, 60.0, myokit.SINGLE_PRECISION, log_interval=0.1)
    rms = np.sqrt(np.mean((P['V'] - O['V']) ** 2))
    ta = activation_times(P['time'], P['V'])
    tb = activation_times(O['time'], O['V'])
    assert np.all(np.isfinite(tb))
    dta = np.max(np.abs(ta - tb))
    print('fp32 rl=%s: V RMS %.3g mV, max activation dt %.3g ms'
          % (rl, rms, dta), flush=True)
    assert rms <= 0.5
    assert dta <= 0.1


def test_fp32_sheet_against_fp64_oracle():
    shape = (48, 32)
    sc = Scenario('lr1991', shape, rl=True, dt=0.01, gx=9.5, gy=9.5,
                  events=[(1, 0, 2, 1000, 0)], paced=(4, shape[1], 0, 0))
    P, O = run_both(sc, 40.0, myokit.SINGLE_PRECISION, log_interval=0.1)
    rms = np.sqrt(np.mean((P['V'] - O['V']) ** 2))
    ta = activation_times(P['time'], P['V'])
    tb = activation_times(O['time'], O['V'])
    ok = np.isfinite(tb)
    assert np.all(np.isfinite(ta[ok]))
    dta = np.max(np.abs(ta[ok] - tb[ok]))
    print('fp32 sheet: V RMS %.3g mV, max activation dt %.3g ms' % (rms, dta),
          flush=True)
    assert rms <= 0.5
    assert dta <= 0.1


def test_log_interval_and_continuation():
    # Log times k * interval (PeriodicTest, test_simulation_log_interval.py)
    # and a run continued into the same log
    s = myokit.SimulationOpenCL(myokit.load_model('example'),
                                myokit.pacing.blocktrain(1000, 0.5), ncells=4,
                                precision=64)
    d = s.run(5, log=['engine.time', 'membrane.V'], log_interval=0.5)
    t = np.asarray(d['engine.time'])
    assert len(t) == 10
    assert np.max(np.abs(t - 0.5 * np.arange(10))) < 1e-9
    d = s.run(2, log=d, log_interval=0.5)
    t = np.asarray(d['engine.time'])
    assert len(t) == 14
    assert np.max(np.abs(t - 0.5 * np.arange(14))) < 1e-9
    assert len(d['3.membrane.V']) == 14


def test_state_and_time_bookkeeping():
    m = myokit.load_model('example')
    s = myokit.SimulationOpenCL(m, None, ncells=(3, 2), precision=64)
    x0 = s.state()
    s.run(1, log=myokit.LOG_NONE)
    assert s.time() == 1
    assert s.state() != x0
    s.reset()
    assert s.time() == 0 and s.state() == x0
    s.pre(2)
    assert s.time() == 0
    assert s.default_state() == s.state() != x0


def test_nan_raises_simulation_error():
    # A huge step makes the model blow up (test_simulation_opencl.py:1493)
    s = myokit.SimulationOpenCL(myokit.load_model('example'),
                                myokit.pacing.blocktrain(1000, 2), ncells=8,
                                precision=32)
    s.set_step_size(5)
    with pytest.raises(myokit.SimulationError):
        s.run(200, log=['engine.time', 'membrane.V'], log_interval=5)


@pytest.mark.parametrize('shape', [(128, 96), (100, 70)])
def test_fp32_persistent_sheet(shape):
    # Large enough for the persistent sheet driver (register-resident tiles
    # with neighbour exchange); the second shape has partial edge tiles
    sc = Scenario('lr1991', shape, rl=True, dt=0.01, gx=9.5, gy=9.5,
                  events=[(1, 0, 2, 1000, 0)], paced=(4, shape[1], 0, 0))
    P, O = run_both(sc, 20.0, myokit.SINGLE_PRECISION, log_interval=0.25)
    assert P['sim'].engine_driver() == 'sheet'
    rms = np.sqrt(np.mean((P['V'] - O['V']) ** 2))
    ta = activation_times(P['time'], P['V'])
    tb = activation_times(O['time'], O['V'])
    ok = np.isfinite(tb)
    assert ok.sum() > 100
    dta = np.max(np.abs(ta[ok] - tb[ok]))
    print('sheet %s: V RMS %.3g mV, max activation dt %.3g ms'
          % (shape, rms, dta), flush=True)
    assert rms <= 0.5
    assert dta <= 0.1
    # Same run on the one-step-per-launch grid driver
    g = sc.product(myokit.SINGLE_PRECISION)
    g.set_engine_option('driver', 2)
    d = g.run(20.0, log=['membrane.V'], log_interval=0.25)
    assert g.engine_driver() == 'grid'
    nx = shape[0]
    Vg = np.stack([np.asarray(d['%d.%d.membrane.V' % (c % nx, c // nx)])
                   for c in range(nx * shape[1])], axis=1)
    dd = np.max(np.abs(Vg - P['V']))
    rg = np.sqrt(np.mean((Vg - P['V']) ** 2))
    print('sheet vs grid max|dV| %.3g mV, RMS %.3g mV' % (dd, rg), flush=True)
    # The same single-precision arithmetic in both kernels (multiply-adds
    # fused within source expressions only): identical results
    assert dd == 0.0


def test_fp32_sheet_continues_across_runs():
    # State stays on the device between runs; two 10 ms runs equal one
    # 20 ms run (bit for bit: single precision fuses multiply-adds only
    # within one source expression, so every launch computes alike)
    shape = (128, 96)
    sc = Scenario('lr1991', shape, rl=True, dt=0.01, gx=9.5, gy=9.5,
                  events=[(1, 0, 2, 1000, 0)], paced=(4, shape[1], 0, 0))
    a = sc.product(32)
    b = sc.product(32)
    a.run(20.0, log=myokit.LOG_NONE)
    b.run(10.0, log=myokit.LOG_NONE)
    b.run(10.0, log=myokit.LOG_NONE)
    assert a.engine_driver() == 'sheet'
    np.testing.assert_array_equal(a.state(), b.state())


def test_fp32_sheet_single_step_launches():
    # A run per time step makes every sheet launch a one-step launch; the
    # result must equal one long run bit for bit (each launch reads its
    # start state, halos included, from a buffer no tile writes)
    shape = (128, 96)
    sc = Scenario('lr1991', shape, rl=True, dt=0.01, gx=9.5, gy=9.5,
                  events=[(1, 0, 2, 1000, 0)], paced=(4, shape[1], 0, 0))
    a = sc.product(32)
    b = sc.product(32)
    a.run(3.0, log=myokit.LOG_NONE)
    b.run(3.0, log=myokit.LOG_NONE)
    a.run(1.0, log=myokit.LOG_NONE)
    for _ in range(100):
        b.run(0.01, log=myokit.LOG_NONE)
    assert a.engine_driver() == b.engine_driver() == 'sheet'
    np.testing.assert_array_equal(a.state(), b.state())


@pytest.mark.parametrize('shape', [(128, 96), (200, 40)])
def test_fp64_persistent_sheet_bit_exact(shape):
    # The sheet driver's neighbour exchange, checked bit for bit in double
    # precision: any stale halo value would show as a difference
    sc = Scenario('lr1991', shape, rl=True, dt=0.01, gx=9.5, gy=7.0,
                  events=[(1, 0, 2, 1000, 0)], paced=(4, 20, 0, 10))
    P, O = run_both(sc, 12.0, myokit.DOUBLE_PRECISION, log_interval=0.5)
    assert P['sim'].engine_driver() == 'sheet'
    assert _report('fp64 sheet ' + str(shape), P, O) <= FP64_TOL
    assert np.max(np.abs(P['state'] - O['state'])) <= FP64_TOL


def test_fp64_graph_cell_order_boundaries():
    # set_connections graphs live on the device in reverse Cuthill-McKee
    # order; every per-cell array is permuted where it crosses the C ABI.
    # A random geometric graph (the order far from the identity) with what
    # crosses that boundary: a paced rectangle (the first 5 cells), a
    # per-cell field, one cell's state set before the run, the state read
    # back after it, and a second graph mid-simulation (the state comes back
    # in the old order and goes up in the new one). Bit for bit against the
    # oracle, which keeps the user's order.
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(12)
    n = 400
    pts = rng.uniform(0, 1, (n, 2))

    def graph(r, seed):
        g = np.random.default_rng(seed)
        pairs = cKDTree(pts).query_pairs(r, output_type='ndarray')
        return [(int(i), int(j), float(g.uniform(5, 12))) for i, j in pairs]
    c1, c2 = graph(0.09, 1), graph(0.07, 2)
    gna = 16.0 * rng.uniform(0.8, 1.2, n)
    sc = Scenario('lr1991', n, dt=0.005, events=[(1, 0, 2, 1000, 0)],
                  paced=(5,), connections=c1, fields={'ina.gNa': gna})
    s = sc.product(myokit.DOUBLE_PRECISION)
    perm, _ = s._cell_order()
    assert perm is not None and not np.array_equal(perm, np.arange(n))
    o = sc.oracle()
    y = np.array(s.state()).reshape(n, -1)
    y[17, 0] = -30.0
    s.set_state(y[17], 17)
    o.state = o.state.reshape(n, -1).copy()
    o.state[17, 0] = -30.0
    o.state = o.state.reshape(-1)
    keys = [str(c) + '.membrane.V' for c in range(n)]
    d = s.run(10.0, log=['membrane.V'], log_interval=0.5)
    r = o.run(10.0, log_vars=['membrane.V'], log_interval=0.5)
    V = np.stack([np.asarray(d[k]) for k in keys], axis=1)
    print('graph order: max|dV| %g, state %g' % (
        np.max(np.abs(V - r['membrane.V'])),
        np.max(np.abs(np.array(s.state()) - o.state))), flush=True)
    np.testing.assert_array_equal(V, r['membrane.V'])
    np.testing.assert_array_equal(np.array(s.state()), o.state)
    # a second graph: new device order, the run continues
    s.run(2.0, log=myokit.LOG_NONE)
    o.run(2.0, log_vars=['membrane.V'], log_interval=100.0)
    s.set_connections(c2)
    o.connections = [(min(i, j), max(i, j), g) for i, j, g in c2]
    d = s.run(6.0, log=['membrane.V'], log_interval=0.5)
    r = o.run(6.0, log_vars=['membrane.V'], log_interval=0.5)
    V = np.stack([np.asarray(d[k]) for k in keys], axis=1)
    np.testing.assert_array_equal(V, r['membrane.V'])
    np.testing.assert_array_equal(np.array(s.state()), o.state)


@pytest.mark.parametrize('driver', ['grid', 'sheet'])
def test_fp64_transposed_sheet(driver):
    # The (nx, ny) sheet with (gx, gy) and a corner stimulus equals the
    # transposed sheet with (gy, gx), bit for bit, on the device too (the
    # CPU test of the same property pins the oracle's y term)
    nx, ny = 80, 64
    out = []
    for shape, g, paced in (((nx, ny), (9.5, 4.0), (6, 5, 0, 0)),
                            ((ny, nx), (4.0, 9.5), (5, 6, 0, 0))):
        s = myokit.SimulationOpenCL(
            myokit.load_model('example'), myokit.pacing.blocktrain(1000, 2),
            ncells=shape, rl=True, precision=myokit.DOUBLE_PRECISION)
        s.set_tuning(driver=driver)
        s.set_step_size(0.01)
        s.set_conductance(*g)
        s.set_paced_cells(*paced)
        s.run(10.0, log=[], log_interval=1.0)
        assert s.engine_driver() == driver
        out.append(np.array(s.state()).reshape(shape[1], shape[0], -1))
    np.testing.assert_array_equal(out[0], out[1].transpose(1, 0, 2))
    V = out[0][:, :, 0]
    assert np.max(V[0, :]) - np.min(V[0, :]) > 10
    assert np.max(V[:, 0]) - np.min(V[:, 0]) > 10
