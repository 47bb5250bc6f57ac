"""
Single-cell accuracy on the GPU against the CVODES stand-in (oracle/ode.py:
scipy Radau at rtol = atol = 1e-10, restarted at every pacing
discontinuity; SUNDIALS is absent here, so CVODES itself is parity
unpinned). The scenarios and bounds are those of the reference's
``myokit/tests/test_simulation_opencl_vs_cvode.py:28-572`` (MRMS as in
Marsh, Ziaratgahi and Spiteri 2012: sqrt(mean(((a - b) / (1 + |b|))^2))),
plus native maths in single precision, and the discrete adjoint of a single
cell against gradients formed from forward sensitivities
(``cvodessim.py:20-164``, ``cmodel.py:244-277``).
"""
import os

import numpy as np
import pytest

import myokit_amd as myokit
from oracle.ode import OdeCell

pytestmark = pytest.mark.gpu

DATA = os.path.join(os.path.dirname(__file__), 'data')
LV = ['engine.time', 'membrane.V', 'engine.pace', 'ica.ICa']


def mrms(a, b):
    r = (np.asarray(a) - np.asarray(b)) / (1 + np.abs(b))
    return float(np.sqrt(np.sum(r ** 2) / len(r)))


def _lr(start):
    m, p, _ = myokit.load(os.path.join(DATA, 'lr-1991.mmt'))
    e = p.head()
    q = myokit.Protocol()
    q.schedule(level=e.level(), duration=e.duration(), period=600,
               start=start)
    return m, q


def _compare(name, start, dt, tmax, rl, native=False,
             precision=myokit.DOUBLE_PRECISION):
    m, p = _lr(start)
    s = myokit.SimulationOpenCL(m, p, ncells=1, rl=rl, native_maths=native,
                                precision=precision)
    s.set_step_size(dt)
    d = s.run(tmax, LV, log_interval=dt)
    r = OdeCell(m, p).run(tmax, log_vars=['membrane.V', 'ica.ICa'],
                          log_interval=dt)
    t = np.asarray(d['engine.time'], dtype=float)
    assert t.shape == r['time'].shape
    e0 = float(np.max(np.abs(t - r['time'])))
    e1 = float(np.sum((np.asarray(d['engine.pace']) - r['pace']) ** 2))
    e2 = mrms(d['0.membrane.V'], r['membrane.V'])
    e3 = mrms(d['0.ica.ICa'], r['ica.ICa'])
    print(name, 'time', e0, 'pace', e1, 'MRMS V', e2, 'MRMS ICa', e3,
          flush=True)
    return e0, e1, e2, e3


def test_event_at_t0():
    e0, e1, e2, e3 = _compare('event at t0', 0, 0.1, 1300, True)
    assert e0 < 1e-14 and e1 < 1e-14
    assert e2 < 0.1 and e3 < 0.05


def test_event_at_step_size_multiple():
    e0, e1, e2, e3 = _compare('event at multiple', 10, 0.1, 1300, True)
    assert e0 < 1e-14 and e1 < 1e-14
    assert e2 < 0.1 and e3 < 0.05


def test_event_not_at_multiple():
    e0, e1, e2, e3 = _compare('event off grid', 1.05, 0.1, 1300, True)
    assert e0 < 1e-14 and e1 < 1e-14
    assert e2 < 0.2 and e3 < 0.05


def test_event_at_step_size_multiple_no_rl():
    e0, e1, e2, e3 = _compare('no RL', 10, 0.01, 200, False)
    assert e2 < 0.05 and e3 < 0.01


@pytest.mark.parametrize('precision', [myokit.DOUBLE_PRECISION,
                                       myokit.SINGLE_PRECISION])
def test_native_maths(precision):
    e0, e1, e2, e3 = _compare('native maths ' + str(precision), 0, 0.005, 10,
                              False, native=True, precision=precision)
    assert e2 < 0.06 and e3 < 0.004


def test_rl():
    e0, e1, e2, e3 = _compare('RL', 0, 0.005, 10, True)
    assert e2 < 0.1 and e3 < 0.006


def _br():
    m = myokit.load_model(os.path.join(DATA, 'beeler-1977-model.mmt'))
    p = myokit.pacing.blocktrain(duration=2, offset=1, period=1000)
    return m, p


def test_fields():
    m, p = _br()
    t, dt = 6.5, 0.1
    sb = myokit.SimulationOpenCL(m, p, ncells=3, diffusion=False)
    sb.set_field('membrane.C', [0.9, 1.1, 1.2])
    sb.set_step_size(0.001)
    dx = sb.run(t, log=['engine.time', 'membrane.V'], log_interval=dt)
    for i, c in enumerate((0.9, 1.1, 1.2)):
        r = OdeCell(m, p, {'membrane.C': c}).run(
            t, log_vars=['membrane.V'], log_interval=dt)
        e = float(np.max(np.abs(r['membrane.V']
                                - np.asarray(dx[str(i) + '.membrane.V']))))
        print('field C =', c, 'max|dV|', e, flush=True)
        assert e < 0.5


def test_set_constant():
    m, p = _br()
    sb = myokit.SimulationOpenCL(m, p, ncells=1)
    sb.set_constant('membrane.C', 1.5)
    sb.set_step_size(0.001)
    d = sb.run(6.5, log=['engine.time', 'membrane.V'], log_interval=0.1)
    m2 = m.clone()
    m2.get('membrane.C').set_rhs(1.5)
    r = OdeCell(m2, p).run(6.5, log_vars=['membrane.V'], log_interval=0.1)
    e = float(np.max(np.abs(r['membrane.V']
                            - np.asarray(d['0.membrane.V']))))
    print('set_constant max|dV|', e, flush=True)
    assert e < 0.5
    # and the constant is in effect (C = 1 gives a different trace)
    m3 = m.clone()
    r3 = OdeCell(m3, p).run(6.5, log_vars=['membrane.V'], log_interval=0.1)
    assert np.max(np.abs(r3['membrane.V'] - r['membrane.V'])) > 1


def test_adjoint_against_forward_sensitivities():
    # L = sum_s (V(t_s) - T)^2 over the N steps of a 20 ms run (T = -80 mV),
    # for the gNa and gCa fields of one cell. The discrete adjoint is the
    # exact gradient of the forward-Euler loss; the forward sensitivities
    # give the gradient of the same sum over the exact solution. They
    # differ by O(dt): the relative difference must be small and halve
    # with dt.
    from myokit_amd.adjoint import TissueAdjoint
    m, p = _lr(1)
    fields = {'ina.gNa': 16.0, 'ica.gCa': 0.09}
    T, dur = -80.0, 20.0
    names = list(fields)
    ode = OdeCell(m, p, fields)
    diffs = []
    for dt in (0.002, 0.001):
        s = myokit.SimulationOpenCL(m, p, ncells=1,
                                    precision=myokit.DOUBLE_PRECISION)
        s.set_step_size(dt)
        for k, v in fields.items():
            s.set_field(k, [v])
        adj = TissueAdjoint(s, [0], checkpoint_every=500)
        n = adj.n_steps(dur)
        loss, grads, _, _ = adj.loss_and_gradient(dur, np.full((n, 1), T))
        ode.state = np.array(ode.model.initial_values(True), dtype=float)
        ode.time = 0.0
        r = ode.run(dur + dt / 2, log_vars=['membrane.V'], log_interval=dt,
                    sensitivities=names)
        assert len(r['time']) == n + 1
        iv = [v.qname() for v in ode.model.states()].index('membrane.V')
        V = r['membrane.V'][1:]
        rel = []
        for k in names:
            g_ode = float(np.sum(2 * (V - T) * r[('dV', k)][1:, iv]))
            g_adj = float(grads[k][0])
            rel.append(abs(g_adj - g_ode) / abs(g_ode))
            print('dt', dt, k, 'adjoint', g_adj, 'sensitivities', g_ode,
                  'rel', rel[-1], flush=True)
        lo = float(np.sum((V - T) ** 2))
        print('dt', dt, 'loss', loss, 'ode', lo, flush=True)
        diffs.append(rel)
    for a, b in zip(diffs[0], diffs[1]):
        assert b < 0.02
        assert 1.5 < a / b < 2.6
