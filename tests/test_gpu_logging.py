"""
GPU parity of everything a log row holds, not only V: the diffusion current
(membrane.i_diff), intermediaries (ica.ICa, isi.Isi), a gate state, the
bound engine.time and engine.pace, and LOG_STATE logs, on every execution
driver (cells, block, grid, sheet) and every coupling (none, stencil,
conductance field, graph), in double precision against the CPU oracle.

The reference writes a row from the state at the row's time t, before the
step from t (openclsim.c:1108-1131, inter_log of openclsim.cl:298-367), and
its own test compares i_diff, Isi and pace of SimulationOpenCL with
Simulation1d (test_simulation_opencl_vs_sim1d.py:118-136). The oracle is
pinned to Simulation1d golden runs for the same quantities
(tests/test_oracle.py); here the device must equal the oracle bit for bit
(same operation order, same portable math).
"""
import os

import numpy as np
import pytest

import myokit_amd as myokit
from helpers import Scenario, load

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), 'golden')
LR_VARS = ['membrane.V', 'membrane.i_diff', 'ica.ICa', 'ina.m']


def _keys(sc, n, var):
    if isinstance(sc.ncells, tuple):
        nx = sc.ncells[0]
        return [str(c % nx) + '.' + str(c // nx) + '.' + var
                for c in range(n)]
    return [str(c) + '.' + var for c in range(n)]


def run_logged(sc, duration, variables, log_interval, driver=None,
               precision=myokit.DOUBLE_PRECISION, log=None):
    """Runs the scenario on the device and on the oracle, logging
    ``variables`` of every cell plus engine.time and engine.pace. Returns
    (product arrays, oracle arrays, driver name)."""
    s = sc.product(precision)
    if driver is not None:
        s.set_tuning(driver=driver)
    if log is None:
        log = ['engine.time', 'engine.pace'] + [
            k for v in variables for k in _keys(sc, s._ntotal, v)]
    d = s.run(duration, log=log, log_interval=log_interval)
    n = s._ntotal
    P = {'time': np.asarray(d['engine.time'], dtype=float),
         'pace': np.asarray(d['engine.pace'], dtype=float),
         'state': np.array(s.state())}
    for v in variables:
        P[v] = np.stack([np.asarray(d[k], dtype=float)
                         for k in _keys(sc, n, v)], axis=1)
    o = sc.oracle()
    r = o.run(duration, log_vars=list(variables), log_interval=log_interval)
    O = {'time': r['time'], 'pace': r['pace'], 'state': o.state}
    for v in variables:
        O[v] = r[v]
    return P, O, s.engine_driver()


def assert_rows_equal(name, P, O, variables):
    assert P['time'].shape == O['time'].shape
    np.testing.assert_array_equal(P['time'], O['time'])
    np.testing.assert_array_equal(P['pace'], O['pace'])
    for v in variables:
        assert P[v].shape == O[v].shape, v
        d = float(np.max(np.abs(P[v] - O[v]))) if P[v].size else 0.0
        print(name, v, 'rows', P[v].shape, 'max|d| =', d, flush=True)
        np.testing.assert_array_equal(P[v], O[v], err_msg=name + ' ' + v)
    np.testing.assert_array_equal(P['state'], O['state'])


def _geometric_graph(n, seed, degree=6.0):
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    pts = rng.uniform(0, 1, size=(n, 2))
    r = np.sqrt(degree / (np.pi * n))
    pairs = np.array(sorted(cKDTree(pts).query_pairs(r)), dtype=np.int64)
    g = rng.uniform(5, 12, size=len(pairs))
    return [(int(i), int(j), float(w)) for (i, j), w in zip(pairs, g)]


@pytest.mark.parametrize('driver', ['block', 'grid'])
def test_logged_quantities_cable(driver):
    sc = Scenario('lr1991', 40, dt=0.005, gx=9.5, paced=(5,))
    P, O, used = run_logged(sc, 12.0, LR_VARS, 0.25, driver)
    assert used == driver
    assert np.max(np.abs(O['membrane.i_diff'])) > 10
    assert np.max(O['pace']) == 1
    assert_rows_equal('cable ' + driver, P, O, LR_VARS)


def test_logged_quantities_cells_with_fields():
    rng = np.random.default_rng(11)
    n = 300
    fields = {'ina.gNa': 16 * np.exp(rng.uniform(np.log(.5), np.log(2), n)),
              'ik.IK.gK': 0.282 * np.exp(rng.uniform(np.log(.5), np.log(2),
                                                     n))}
    sc = Scenario('lr1991', n, diffusion=False, dt=0.005, fields=fields)
    # without diffusion, membrane.i_diff is an unbound constant, which a log
    # cannot hold (as in the reference)
    variables = ['membrane.V', 'ica.ICa', 'ina.m', 'ina.INa']
    P, O, used = run_logged(sc, 8.0, variables, 0.5)
    assert used == 'cells'
    assert_rows_equal('cells', P, O, variables)


@pytest.mark.parametrize('driver,shape', [('block', (24, 16)),
                                          ('grid', (48, 40)),
                                          ('sheet', (80, 64))])
def test_logged_quantities_sheet_rl(driver, shape):
    # A corner stimulus, so that the x and y terms of the stencil differ
    # (gx != gy) and both are present in every logged i_diff
    sc = Scenario('lr1991', shape, rl=True, dt=0.01, gx=9.5, gy=6.0,
                  paced=(5, 4, 0, 0))
    P, O, used = run_logged(sc, 8.0, LR_VARS, 0.5, driver)
    assert used == driver
    assert_rows_equal('sheet ' + driver + ' ' + str(shape), P, O, LR_VARS)


@pytest.mark.parametrize('driver', ['block', 'grid'])
def test_logged_quantities_conductance_field(driver):
    rng = np.random.default_rng(5)
    nx, ny = 20, 12
    gx = rng.uniform(5, 12, size=(ny, nx - 1))
    gy = rng.uniform(5, 12, size=(ny - 1, nx))
    sc = Scenario('lr1991', (nx, ny), rl=True, dt=0.01, paced=(3, 3, 0, 0),
                  conductance_field=(gx, gy))
    P, O, used = run_logged(sc, 8.0, LR_VARS, 0.5, driver)
    assert used == driver
    assert_rows_equal('hetero ' + driver, P, O, LR_VARS)


@pytest.mark.parametrize('driver', ['block', 'grid'])
def test_logged_quantities_graph(driver):
    # A graph the Python class renumbers (reverse Cuthill-McKee) on the
    # device: every logged column must come back in the caller's order
    n = 400
    conns = _geometric_graph(n, 3)
    sc = Scenario('lr1991', n, rl=True, dt=0.01, paced_list=list(range(8)),
                  connections=conns)
    P, O, used = run_logged(sc, 8.0, LR_VARS, 0.5, driver)
    assert used == driver
    assert_rows_equal('graph ' + driver, P, O, LR_VARS)


@pytest.mark.parametrize('same_g', [True, False])
def test_logged_quantities_graph_copies(same_g):
    # Four copies of a 2,500-cell graph (a batch of parameter sets, with
    # their own fields), which the device holds with the copies
    # interleaved; with the same conductances in every copy the library
    # keeps one copy's connection rows for all of them. Every logged column
    # in the caller's order, bit for bit
    m, K = 2500, 4
    base = _geometric_graph(m, 9)
    rng = np.random.default_rng(12)
    conns = [(i + k * m, j + k * m,
              g if same_g else float(rng.uniform(5, 12)))
             for k in range(K) for (i, j, g) in base]
    fields = {'ina.gNa': 16 * np.exp(rng.uniform(np.log(.5), np.log(2),
                                                  m * K))}
    paced = [c + k * m for k in range(K) for c in range(8)]
    sc = Scenario('lr1991', m * K, rl=True, dt=0.01, paced_list=paced,
                  connections=conns, fields=fields)
    s = sc.product(myokit.DOUBLE_PRECISION)
    perm, iperm = s._cell_order()
    assert perm is not None and perm[1] - perm[0] == m  # interleaved
    assert s._copies == K
    P, O, used = run_logged(sc, 6.0, LR_VARS, 0.5)
    assert used == 'grid'
    assert_rows_equal('graph copies ' + used, P, O, LR_VARS)


def test_log_state_flag():
    # LOG_STATE + LOG_BOUND: every state of every cell and the bound
    # variables, keyed as the reference keys them
    sc = Scenario('lr1991', 12, dt=0.005, gx=9.5, paced=(3,))
    m = load('lr1991')
    states = [v.qname() for v in m.states()]
    s = sc.product(myokit.DOUBLE_PRECISION)
    d = s.run(6.0, log=myokit.LOG_STATE + myokit.LOG_BOUND,
              log_interval=0.25)
    assert 'engine.time' in d and 'engine.pace' in d
    o = sc.oracle()
    r = o.run(6.0, log_vars=states, log_interval=0.25)
    np.testing.assert_array_equal(np.asarray(d['engine.time']), r['time'])
    np.testing.assert_array_equal(np.asarray(d['engine.pace']), r['pace'])
    for v in states:
        got = np.stack([np.asarray(d[str(c) + '.' + v], dtype=float)
                        for c in range(12)], axis=1)
        np.testing.assert_array_equal(got, r[v], err_msg=v)
    print('LOG_STATE', len(d.keys()), 'columns equal', flush=True)


@pytest.mark.parametrize('name', ['sim1d_br1977_reftest', 'sim1d_lr1991_rl'])
def test_logged_quantities_against_simulation1d(name):
    # The reference's own comparison (test_simulation_opencl_vs_sim1d.py:
    # 28-136): time and pace within 1e-17, V, i_diff and the intermediary
    # within 1e-13 of the reference's Simulation1d
    g = np.load(os.path.join(GOLD, name + '.npz'))
    n, npaced, gx, rl, dt, dur, li = g['config']
    model = 'br1977' if 'br1977' in name else 'lr1991'
    m = load(model)
    if np.isfinite(g['amplitude']):
        m.get('membrane.i_stim.amplitude').set_rhs(float(g['amplitude']))
    p = myokit.Protocol()
    p.schedule(*g['event'])
    s = myokit.SimulationOpenCL(m, p, ncells=int(n), rl=bool(rl),
                                precision=myokit.DOUBLE_PRECISION)
    s.set_conductance(gx)
    s.set_paced_cells(int(npaced))
    s.set_step_size(dt)
    inter = str(g['inter_name'])
    lv = ['engine.time', 'membrane.V', 'engine.pace', inter,
          'membrane.i_diff']
    d = s.run(dur, log=lv, log_interval=li)
    assert np.max(np.abs(np.asarray(d['engine.time']) - g['time'])) < 1e-17
    assert np.max(np.abs(np.asarray(d['engine.pace']) - g['pace'])) < 1e-17
    for var, key in (('membrane.V', 'V'), ('membrane.i_diff', 'i_diff'),
                     (inter, 'inter')):
        got = np.stack([np.asarray(d[str(c) + '.' + var], dtype=float)
                        for c in range(int(n))], axis=1)
        e = float(np.max(np.abs(got - g[key])))
        print(name, var, 'max|d| vs Simulation1d =', e, flush=True)
        assert e < 1e-11
