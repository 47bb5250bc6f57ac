"""
CPU tests that pin the single-cell ODE solutions standing in for CVODES
(oracle/ode.py; SUNDIALS is absent here, so CVODES itself is parity
unpinned): forward Euler from the tissue oracle converges to them at first
order, a linear model's closed form is reproduced, and the forward
sensitivities equal central differences of whole runs.
"""
import os

import numpy as np

import myokit_amd as myokit
from oracle.cgen import OracleSimulation
from oracle.ode import OdeCell

DATA = os.path.join(os.path.dirname(__file__), 'data')


def _lr():
    m, p, _ = myokit.load(os.path.join(DATA, 'lr-1991.mmt'))
    e = p.head()
    q = myokit.Protocol()
    q.schedule(level=e.level(), duration=e.duration(), period=600, start=1)
    return m, q


def test_forward_euler_converges_to_the_ode_solution():
    # The tissue oracle's single-cell forward Euler (the reference's
    # cell_step) against the Radau solution: errors halve with dt (first
    # order), down to well under a millivolt
    m, p = _lr()
    r = OdeCell(m, p).run(40.0, log_vars=['membrane.V'], log_interval=0.5)
    errs = []
    for dt in (0.004, 0.002, 0.001):
        o = OracleSimulation(m, p, ncells=1, diffusion=False)
        o.dt = dt
        ro = o.run(40.0, log_vars=['membrane.V'], log_interval=0.5)
        np.testing.assert_array_equal(ro['time'], r['time'])
        np.testing.assert_array_equal(ro['pace'], r['pace'])
        errs.append(np.max(np.abs(ro['membrane.V'][:, 0] - r['membrane.V'])))
    print('FE error at dt 0.004, 0.002, 0.001:', errs)
    assert errs[2] < 0.2
    for a, b in zip(errs, errs[1:]):
        assert 1.7 < a / b < 2.3
    assert np.max(r['membrane.V']) > 30


def test_linear_model_closed_form():
    # dx/dt = -k x + pace: piecewise exponentials, through two pacing
    # discontinuities
    m = myokit.parse_model('''
[[model]]
c.x = 1
[engine]
time = 0 bind time
pace = 0 bind pace
[c]
k = 0.3
dot(x) = -k * x + engine.pace
''')
    p = myokit.Protocol()
    p.schedule(level=2, start=1.5, duration=2)
    r = OdeCell(m, p).run(6.0, log_vars=['c.x'], log_interval=0.25)
    k = 0.3
    t = r['time']
    x = np.where(t < 1.5, np.exp(-k * t), 0)
    x15 = np.exp(-k * 1.5)
    on = (t >= 1.5) & (t < 3.5)
    x = np.where(on, 2 / k + (x15 - 2 / k) * np.exp(-k * (t - 1.5)), x)
    x35 = 2 / k + (x15 - 2 / k) * np.exp(-k * 2)
    x = np.where(t >= 3.5, x35 * np.exp(-k * (t - 3.5)), x)
    assert np.max(np.abs(r['c.x'] - x)) < 1e-9
    np.testing.assert_array_equal(r['pace'], np.where(on, 2.0, 0.0))


def test_sensitivities_equal_central_differences_of_runs():
    m, p = _lr()
    states = [v.qname() for v in m.states()]
    iv = states.index('membrane.V')
    fields = {'ina.gNa': 16.0, 'ica.gCa': 0.09}
    base = OdeCell(m, p, fields).run(
        30.0, log_vars=['membrane.V'], log_interval=0.5,
        sensitivities=['ina.gNa', 'ica.gCa'])
    for name, h in (('ina.gNa', 1e-3), ('ica.gCa', 1e-5)):
        up = dict(fields)
        up[name] += h
        dn = dict(fields)
        dn[name] -= h
        a = OdeCell(m, p, up).run(30.0, log_vars=['membrane.V'],
                                  log_interval=0.5)['membrane.V']
        b = OdeCell(m, p, dn).run(30.0, log_vars=['membrane.V'],
                                  log_interval=0.5)['membrane.V']
        fd = (a - b) / (2 * h)
        s = base[('dV', name)][:, iv]
        err = np.max(np.abs(s - fd)) / np.max(np.abs(fd))
        print(name, 'sensitivity vs central difference of runs:', err)
        assert err < 1e-4
        assert np.max(np.abs(fd)) > 0
