"""
GPU tests of the chain driver (``mk_chain``): graphs whose connections form
disjoint chains (cables), held as one-wave tiles laid end to end, with cables
longer than a tile exchanging edge potentials between tiles. Double
precision must equal the oracle bit for bit (the diffusion current sums a
cell's connection terms in its connection-list order, as the reference's
edge scatter does when its atomics run in list order, openclsim.cl:537-556);
single precision must equal the block driver's single precision bit for bit
and the double-precision oracle within the tolerances.
"""
import numpy as np
import pytest

import myokit_amd as myokit
from helpers import Scenario, run_both, run_product

pytestmark = pytest.mark.gpu


def _cables(rng, ncables, lo, hi, shuffle=False, flip=False, relabel=False):
    """Connections of ``ncables`` chains with lengths in [lo, hi], the first
    3 cells of each paced; optionally listed in random order, with (j, i)
    instead of (i, j), and with the cells numbered at random."""
    lengths = rng.integers(lo, hi + 1, ncables)
    n = int(lengths.sum())
    conns, paced, off = [], [], 0
    for L in lengths:
        conns += [(off + i, off + i + 1, float(rng.uniform(5, 12)))
                  for i in range(L - 1)]
        paced += list(range(off, off + 3))
        off += L
    if relabel:
        lab = rng.permutation(n)
        conns = [(int(lab[i]), int(lab[j]), g) for i, j, g in conns]
        paced = [int(lab[c]) for c in paced]
    if flip:
        conns = [(j, i, g) if rng.uniform() < 0.5 else (i, j, g)
                 for i, j, g in conns]
    if shuffle:
        conns = [conns[k] for k in rng.permutation(len(conns))]
    return n, conns, paced, lengths


@pytest.mark.parametrize('case', ['ordered', 'shuffled', 'relabelled'])
def test_fp64_chain_against_oracle(case):
    # 40 cables of 20-300 cells: most span several 64-cell tiles, and the
    # tile boundaries fall inside cables as well as between them
    rng = np.random.default_rng({'ordered': 1, 'shuffled': 2,
                                 'relabelled': 3}[case])
    n, conns, paced, lengths = _cables(
        rng, 40, 20, 300, shuffle=case != 'ordered', flip=case != 'ordered',
        relabel=case == 'relabelled')
    sc = Scenario('lr1991', n, dt=0.005, connections=conns,
                  paced_list=paced, events=[(1, 0, 2, 1000, 0)])
    P, O = run_both(sc, 20.0, myokit.DOUBLE_PRECISION, log_interval=0.5)
    assert P['sim'].engine_driver() == 'chain'
    d = float(np.max(np.abs(P['V'] - O['V'])))
    print('chain fp64', case, n, 'cells: max|dV|', d, flush=True)
    np.testing.assert_array_equal(P['V'], O['V'])
    np.testing.assert_array_equal(P['state'], O['state'])
    assert np.max(O['V']) > 0


def test_fp32_chain_equals_block_and_oracle():
    rng = np.random.default_rng(4)
    n, conns, paced, lengths = _cables(rng, 24, 30, 400, shuffle=True,
                                       flip=True)
    sc = Scenario('lr1991', n, dt=0.005, connections=conns,
                  paced_list=paced, events=[(1, 0, 2, 1000, 0)])
    P, O = run_both(sc, 30.0, myokit.SINGLE_PRECISION, log_interval=0.5)
    assert P['sim'].engine_driver() == 'chain'
    s = sc.product(myokit.SINGLE_PRECISION)
    s.set_tuning(driver='block')
    d = s.run(30.0, log=['engine.time'] + [str(c) + '.membrane.V'
                                             for c in range(n)],
              log_interval=0.5)
    assert s.engine_driver() == 'block'
    B = np.stack([np.asarray(d[str(c) + '.membrane.V']) for c in range(n)],
                 axis=1)
    np.testing.assert_array_equal(P['V'], B)
    rms = float(np.sqrt(np.mean((P['V'] - O['V']) ** 2)))
    print('chain fp32 V RMS vs fp64 oracle', rms, flush=True)
    assert rms <= 0.5


def test_chain_launch_splits_and_logged_quantities():
    # Runs split into many launches (a log row every 0.05 ms) carry the
    # tiles' step counts across launches; intermediaries, i_diff and a gate
    # logged from the chain driver equal the oracle's
    rng = np.random.default_rng(5)
    n, conns, paced, lengths = _cables(rng, 6, 100, 200, shuffle=True)
    sc = Scenario('lr1991', n, dt=0.005, connections=conns,
                  paced_list=paced, events=[(1, 0, 2, 1000, 0)])
    variables = ['membrane.V', 'membrane.i_diff', 'ica.ICa', 'ina.m']
    s = sc.product(myokit.DOUBLE_PRECISION)
    keys = ['engine.time', 'engine.pace'] + [
        str(c) + '.' + v for v in variables for c in range(n)]
    d = s.run(6.0, log=keys, log_interval=0.05)
    assert s.engine_driver() == 'chain'
    r = sc.oracle().run(6.0, log_vars=variables, log_interval=0.05)
    np.testing.assert_array_equal(np.asarray(d['engine.time']), r['time'])
    np.testing.assert_array_equal(np.asarray(d['engine.pace']), r['pace'])
    for v in variables:
        got = np.stack([np.asarray(d[str(c) + '.' + v]) for c in range(n)],
                       axis=1)
        np.testing.assert_array_equal(got, r[v], err_msg=v)


def test_chain_adjoint_runs_on_the_grid_layout():
    # The adjoint of a chain-shaped graph: equal to the adjoint with the
    # grid driver forced (the chain layout's adjoint runs there)
    from myokit_amd.adjoint import TissueAdjoint
    rng = np.random.default_rng(6)
    n, conns, paced, lengths = _cables(rng, 8, 40, 120)
    out = []
    for driver in (None, 'grid'):
        s = myokit.SimulationOpenCL(
            myokit.load_model('example'), myokit.pacing.blocktrain(1000, 2),
            ncells=n, precision=myokit.DOUBLE_PRECISION)
        s.set_step_size(0.005)
        s.set_connections(conns)
        s.set_paced_cell_list(paced)
        s.set_field('ina.gNa', np.full(n, 16.0))
        if driver:
            s.set_tuning(driver=driver)
        adj = TissueAdjoint(s, [0, n // 2, n - 1], checkpoint_every=100)
        target = np.full((adj.n_steps(5.0), 3), -60.0)
        out.append(adj.loss_and_gradient(5.0, target))
        s.run(0.5, log=[])
        print('adjoint with driver', driver, '->', s.engine_driver(),
              flush=True)
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1]['ina.gNa'], out[1][1]['ina.gNa'])
    assert out[0][2] == out[1][2]


def test_chain_throughput_beside_block():
    # The 256 cables of 50-2000 cells of the batching configuration, 10 ms
    # in single precision: the chain driver against one workgroup per cable
    import time
    rng = np.random.default_rng(11)
    lengths = rng.integers(50, 2001, 256)
    n = int(lengths.sum())
    rows, paced, off = [], [], 0
    for L in lengths:
        i = np.arange(off, off + L - 1)
        rows.append(np.stack([i, i + 1, np.full(L - 1, 9.5)], axis=1))
        paced += list(range(off, off + 5))
        off += L
    rows = np.concatenate(rows)
    rates = {}
    for driver in ('chain', 'block'):
        s = myokit.SimulationOpenCL(
            myokit.load_model('example'), myokit.pacing.blocktrain(1000, 2),
            ncells=n, precision=myokit.SINGLE_PRECISION)
        s.set_step_size(0.005)
        s.set_connections(rows)
        s.set_paced_cell_list(paced)
        s.set_tuning(driver=driver)
        s.run_block(1.0, ['membrane.V'], log_interval=0.5)
        t0 = time.perf_counter()
        b = s.run_block(10.0, ['membrane.V'], log_interval=0.5)
        el = time.perf_counter() - t0
        assert s.engine_driver() == driver
        rates[driver] = 2000 * n / el
        assert np.all(np.isfinite(b['membrane.V']))
    print('cables 262k cells: chain %.3g, block %.3g cell-steps/s'
          % (rates['chain'], rates['block']), flush=True)
    assert rates['chain'] > rates['block']
