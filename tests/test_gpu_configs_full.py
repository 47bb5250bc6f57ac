"""
The benchmark configurations at their full sizes, on the GPU, against the
oracle where it finishes in seconds and otherwise double precision against
single precision on the device:

* configuration 5: one parameter set's 262,144-cell geometric graph, 10 ms,
  double precision bit for bit against the oracle and single precision
  within the tolerances; the whole batch (32 sets, 8,388,608 cells) forward
  and adjoint over 2 ms, single-precision gradients against double;
* configuration 3(ii): the 256 cables of 50-2000 cells, 5 ms against the
  oracle (the chain driver; in double precision more tiles than can be
  co-resident, so tiles of more than 64 cells);
* configuration 2: the 1024 x 1024 sheet for the full 2000 ms with S1 and
  the S2 stimulus on the lower-left quadrant at 320 ms (set between two
  runs, as a user of the reference does), single against double precision;
  and the S1-S2 switch on a 96 x 96 sheet against the oracle;
* single-precision conductance fields (block and grid drivers) and native
  maths in single precision against the double-precision oracle.

Bars: double precision max|dV| 0 against the oracle (requirement 1e-6 mV);
single precision V RMS <= 0.5 mV and activation times within 0.1 ms of
double precision.
"""
import time

import numpy as np
import pytest

import myokit_amd as myokit
from helpers import Scenario, activation_times
from oracle.cgen import OracleSimulation

pytestmark = pytest.mark.gpu


def _scales(rng, n):
    return np.exp(rng.uniform(np.log(0.5), np.log(2.0), n))


def _graph(n=262144, seed=21):
    """Configuration 5's seeded random geometric graph (mean degree 6,
    g uniform in [5, 12]) as an (E, 3) array, and the cells within 0.05 of
    the left edge."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    pts = rng.uniform(0, 1, (n, 2))
    r = np.sqrt(6.0 / (np.pi * n))
    pairs = cKDTree(pts).query_pairs(r, output_type='ndarray')
    g = rng.uniform(5, 12, len(pairs))
    rows = np.stack([pairs[:, 0], pairs[:, 1], g], axis=1).astype(float)
    return rows, np.nonzero(pts[:, 0] < 0.05)[0], rng


def _fp32_report(name, t, P, O, t_act=True):
    dV = P - O
    rms = float(np.sqrt(np.mean(dV ** 2)))
    line = '%s fp32: V RMS %.3g mV, max|dV| %.3g mV' % (
        name, rms, float(np.max(np.abs(dV))))
    dact = 0.0
    if t_act:
        a = activation_times(t, P)
        b = activation_times(t, O)
        both = np.isfinite(a) & np.isfinite(b)
        assert np.array_equal(np.isfinite(a), np.isfinite(b))
        dact = float(np.max(np.abs(a[both] - b[both]))) if both.any() else 0.0
        line += ', activation max|dt| %.3g ms over %d cells' % (
            dact, int(both.sum()))
    print(line, flush=True)
    return rms, dact


def test_config5_one_set_forward_full_size():
    """One parameter set of configuration 5 (262,144 cells, the graph held
    in reverse Cuthill-McKee order on the device), RL, dt 0.01 ms, 10 ms,
    V of every cell every 0.1 ms."""
    rows, left, rng = _graph()
    n = 262144
    m = myokit.load_model('example')
    fields = {k: m.get(k).rhs().eval() * _scales(rng, n)
              for k in ('ina.gNa', 'ica.gCa', 'ik.IK.gK')}
    o = OracleSimulation(m, myokit.pacing.blocktrain(1000, 2), ncells=n,
                         rl=True)
    o.dt = 0.01
    o.connections = rows
    o.paced = [int(c) for c in left]
    for k, v in fields.items():
        o.fields[k] = v
    t0 = time.time()
    r = o.run(10.0, log_vars=['membrane.V'], log_interval=0.1)
    print('config5 oracle: %.1f s' % (time.time() - t0), flush=True)
    for prec in (64, 32):
        s = myokit.SimulationOpenCL(m, myokit.pacing.blocktrain(1000, 2),
                                    ncells=n, precision=prec, rl=True)
        s.set_step_size(0.01)
        s.set_connections(rows)
        s.set_paced_cell_list(left.tolist())
        for k, v in fields.items():
            s.set_field(k, v)
        b = s.run_block(10.0, ['membrane.V'], log_interval=0.1)
        V = b['membrane.V'].reshape(len(b['time']), n)
        assert V.shape == r['membrane.V'].shape
        # log times are the engine time cast to the simulation's precision
        np.testing.assert_array_equal(
            b['time'], r['time'].astype(np.float32 if prec == 32 else float))
        if prec == 64:
            d = float(np.max(np.abs(V - r['membrane.V'])))
            print('config5 one set fp64 (%s driver): max|dV| %.3g mV'
                  % (s.engine_driver(), d), flush=True)
            assert d == 0
        else:
            rms, dact = _fp32_report('config5 one set', b['time'], V,
                                     r['membrane.V'])
            assert rms <= 0.5 and dact <= 0.1
    assert np.sum(np.max(r['membrane.V'], axis=0) > 0) > 1000


def test_config5_batch_gradients_full_size():
    """The whole batch of configuration 5 (32 sets, 8,388,608 cells) over
    a 2 ms window (200 steps, checkpoints every 50): single-precision state
    with the default single-precision VJP against the same adjoint in double
    precision. Probes: 16 cells per set; target -60 mV."""
    rows, left, rng = _graph()
    n, sets = 262144, 32
    allrows = np.concatenate([rows + np.array([k * n, k * n, 0.0])
                              for k in range(sets)])
    paced = np.concatenate([left + k * n for k in range(sets)]).tolist()
    m = myokit.load_model('example')
    fields = {k: m.get(k).rhs().eval() * _scales(rng, n * sets)
              for k in ('ina.gNa', 'ica.gCa', 'ik.IK.gK')}
    probes = [k * n + i for k in range(sets) for i in left[:16]]
    from myokit_amd.adjoint import TissueAdjoint
    out = {}
    for prec in (64, 32):
        t0 = time.time()
        s = myokit.SimulationOpenCL(m, myokit.pacing.blocktrain(1000, 2),
                                    ncells=n * sets, precision=prec, rl=True)
        s.set_step_size(0.01)
        s.set_connections(allrows)
        s.set_paced_cell_list(paced)
        for k, v in fields.items():
            s.set_field(k, v)
        adj = TissueAdjoint(s, probes, checkpoint_every=50)
        target = np.full((adj.n_steps(2.0), len(probes)), -60.0)
        out[prec] = adj.loss_and_gradient(2.0, target)
        print('config5 batch fp%d: loss %.10g, %.1f s (phases %s)'
              % (prec, out[prec][0], time.time() - t0, adj.timing()),
              flush=True)
        adj.release()
        del s, adj
    l64, g64, s64, _ = out[64]
    l32, g32, s32, _ = out[32]
    errs = {'loss': abs(l32 - l64) / abs(l64),
            'scale': abs(s32 - s64) / abs(s64)}
    for k in g64:
        errs[k] = float(np.linalg.norm(g32[k] - g64[k])
                        / np.linalg.norm(g64[k]))
    print('config5 batch fp32 vs fp64 relative errors:', errs, flush=True)
    assert max(errs.values()) <= 1e-3


def _cables(seed=11, ncables=256):
    rng = np.random.default_rng(seed)
    lengths = rng.integers(50, 2001, ncables)
    rows, paced, off = [], [], 0
    for L in lengths:
        i = np.arange(off, off + L - 1)
        rows.append(np.stack([i, i + 1, np.full(L - 1, 9.5)], axis=1))
        paced += list(range(off, off + 5))
        off += L
    return int(off), np.concatenate(rows).astype(float), paced, lengths


def test_config3_uneven_cables_full_size():
    """Configuration 3(ii) in full: 256 cables of 50-2000 cells (seeded),
    chains of g = 9.5, the first 5 cells of each paced, FE, dt 0.005 ms,
    5 ms, V of every cell every 0.1 ms."""
    n, rows, paced, lengths = _cables()
    assert lengths.max() > 1024
    m = myokit.load_model('example')
    p = myokit.pacing.blocktrain(1000, 2)
    o = OracleSimulation(m, p, ncells=n)
    o.dt = 0.005
    o.connections = rows
    o.paced = paced
    t0 = time.time()
    r = o.run(5.0, log_vars=['membrane.V'], log_interval=0.1)
    print('config3 cables oracle: %d cells, %.1f s' % (n, time.time() - t0),
          flush=True)
    for prec in (64, 32):
        s = myokit.SimulationOpenCL(m, p, ncells=n, precision=prec)
        s.set_step_size(0.005)
        s.set_connections(rows)
        s.set_paced_cell_list(paced)
        b = s.run_block(5.0, ['membrane.V'], log_interval=0.1)
        assert s.engine_driver() == 'chain'
        V = b['membrane.V'].reshape(len(b['time']), n)
        np.testing.assert_array_equal(
            b['time'], r['time'].astype(np.float32 if prec == 32 else float))
        if prec == 64:
            d = float(np.max(np.abs(V - r['membrane.V'])))
            print('config3 cables fp64: max|dV| %.3g mV' % d, flush=True)
            assert d == 0
        else:
            rms, dact = _fp32_report('config3 cables', b['time'], V,
                                     r['membrane.V'])
            assert rms <= 0.5 and dact <= 0.1


def _s1s2(shape, precision, s2_time, rl=True):
    m = myokit.load_model('example')
    p = myokit.Protocol()
    p.schedule(1, 0, 2)
    p.schedule(1, s2_time, 2)
    s = myokit.SimulationOpenCL(m, p, ncells=shape, rl=rl,
                                precision=precision)
    s.set_step_size(0.01)
    s.set_conductance(9.5, 9.5)
    s.set_paced_cells(4, shape[1], 0, 0)
    return s, m, p


def test_config2_s2_switch_against_oracle():
    """The S1-S2 protocol of configuration 2 on a 96 x 96 sheet (persistent
    sheet driver): run to the S2 time, pace the lower-left quadrant, run
    on; double precision bit for bit against the oracle doing the same."""
    shape, ts2 = (96, 96), 400.0
    s, m, p = _s1s2(shape, 64, ts2)
    o = OracleSimulation(m, p, ncells=shape, rl=True)
    o.dt = 0.01
    o.gx = o.gy = 9.5
    o.set_paced_cells(4, shape[1], 0, 0)
    outs = []
    for seg, quadrant in ((ts2, False), (40.0, True)):
        if quadrant:
            s.set_paced_cells(48, 48, 0, 0)
            o.set_paced_cells(48, 48, 0, 0)
        b = s.run_block(seg, ['membrane.V'], log_interval=1.0)
        r = o.run(seg, log_vars=['membrane.V'], log_interval=1.0)
        V = b['membrane.V'].reshape(len(b['time']), -1)
        np.testing.assert_array_equal(b['time'], r['time'])
        d = float(np.max(np.abs(V - r['membrane.V'])))
        print('config2 S1-S2 96^2 fp64 (%s driver) to t=%g: max|dV| %.3g'
              % (s.engine_driver(), s.time(), d), flush=True)
        assert d == 0
        outs.append(V)
    # S2 excites the quadrant (paced, recovered by then: the action
    # potential lasts about 350 ms)
    assert np.max(outs[1][:5, 0]) > 0


def _config2_run(s2_time, k_max=8):
    """Configuration 2 at 1024 x 1024 in both precisions on the device, S2
    on the lower-left quadrant at ``s2_time`` (set between two runs), 2000
    ms, V of every cell every 2 ms. Returns (V RMS of single against double
    precision, activation times (k_max, n) per precision: the k-th upward
    crossing of -30 mV of each cell, NaN where there is none)."""
    shape = (1024, 1024)
    sims = {prec: _s1s2(shape, prec, s2_time)[0] for prec in (64, 32)}
    n = shape[0] * shape[1]
    sq = 0.0
    count = 0
    prev = {}
    acts = {prec: np.full((k_max, n), np.nan) for prec in sims}
    seen = {prec: np.zeros(n, dtype=np.int64) for prec in sims}
    t_last = None
    segs, t = [], 0.0
    for stop in sorted({s2_time, 2000.0} | set(np.arange(100.0, 2000.0,
                                                         100.0))):
        segs.append(stop - t)
        t = stop
    for seg in segs:
        if abs(sims[64].time() - s2_time) < 1e-9:
            for s in sims.values():
                s.set_paced_cells(512, 512, 0, 0)
        blocks = {}
        for prec, s in sims.items():
            b = s.run_block(seg, ['membrane.V'], log_interval=2.0)
            blocks[prec] = (b['time'], b['membrane.V'].reshape(
                len(b['time']), n))
        t = blocks[64][0]
        np.testing.assert_array_equal(t, blocks[32][0])
        d = blocks[32][1].astype(np.float64) - blocks[64][1]
        sq += float(np.sum(d * d))
        count += d.size
        del d
        for prec in sims:
            V = blocks[prec][1].astype(np.float64)
            tt = t
            if prec in prev:
                V = np.concatenate([prev[prec][None, :], V])
                tt = np.concatenate([[t_last], t])
            up = (V[1:] >= -30) & (V[:-1] < -30)
            rows_, cells = np.nonzero(up)
            v0, v1 = V[rows_, cells], V[rows_ + 1, cells]
            ta = tt[rows_] + (-30 - v0) / (v1 - v0) * (tt[rows_ + 1]
                                                       - tt[rows_])
            order = np.lexsort((ta, cells))
            cs, ts = cells[order], ta[order]
            start = np.ones(len(cs), dtype=bool)
            start[1:] = cs[1:] != cs[:-1]
            first_idx = np.maximum.accumulate(
                np.where(start, np.arange(len(cs)), 0))
            k = seen[prec][cs] + np.arange(len(cs)) - first_idx
            keep = k < k_max
            acts[prec][k[keep], cs[keep]] = ts[keep]
            seen[prec] += np.bincount(cells, minlength=n)
            prev[prec] = blocks[prec][1][-1].astype(np.float64)
        t_last = t[-1]
    return float(np.sqrt(sq / count)), acts, seen


def _compare_activations(acts, seen):
    """Statistics of single against double precision activation times:
    ``dact`` the largest |dt| over the activations of cells that activate
    equally often in both, ``p999`` its 99.9th percentile, ``over`` how many
    of them differ by more than 0.1 ms, ``differ`` the cells whose counts
    differ, ``later`` the activations after the first (double precision)."""
    a, b = acts[32], acts[64]
    same = seen[32] == seen[64]
    both = np.isfinite(a) & np.isfinite(b) & same[None, :]
    d = np.abs(a[both] - b[both])
    f = np.isfinite(a[0]) & np.isfinite(b[0])
    return {'first': float(np.max(np.abs(a[0][f] - b[0][f]))),
            'dact': float(d.max()) if d.size else 0.0,
            'p999': float(np.percentile(d, 99.9)) if d.size else 0.0,
            'over': int(np.sum(d > 0.1)), 'n': int(d.size),
            'differ': int(np.sum(~same)),
            'later': int(np.sum(np.maximum(seen[64] - 1, 0))),
            'most': int(seen[64].max())}


def test_config2_full_s1_s2_fp32_against_fp64():
    """Configuration 2 as specified: 1024 x 1024, RL, dt 0.01 ms, S1 on the
    left 4 columns at 0 ms, S2 on the lower-left quadrant at 320 ms (set
    between two runs), 2000 ms, V of every cell every 2 ms. Single
    precision against double precision on the device (the oracle would
    need hours; the double-precision device path equals the oracle bit for
    bit on the windows above). Reports V RMS over the whole run and the
    activation times of every cell.

    With the bundled LR1991 the S1 wave crosses 100 cells in 17 ms and a
    cell repolarises below -70 mV about 390 ms after its upstroke, so at
    320 ms the whole quadrant is still refractory: S2 excites nothing, and
    every cell activates once (checked in both precisions)."""
    t0 = time.time()
    rms, acts, seen = _config2_run(320.0)
    c = _compare_activations(acts, seen)
    print('config2 1024^2 S1 + S2 at 320 ms, 2000 ms (%.0f s): fp32 vs fp64 '
          'V RMS %.3g mV; activation max|dt| %.3g ms over %d activations; '
          'cells with different activation counts %d; activations after the '
          'first (fp64) %d' % (time.time() - t0, rms, c['dact'], c['n'],
                               c['differ'], c['later']), flush=True)
    assert c['differ'] == 0 and c['later'] == 0
    assert np.all(seen[64] == 1)
    assert c['dact'] <= 0.1
    assert rms <= 0.5


def test_config2_reentry_fp32_against_fp64():
    """The same sheet with S2 at 440 ms, when the S1 wave's back (V below
    -70 mV) has passed about a third of the quadrant: S2 excites the
    recovered part of the quadrant, which propagates into the recovered
    tissue above it while the still refractory part to its right blocks
    it, and re-entry follows (about a million activations after the
    first). The regime where single precision is hardest to keep beside
    double precision; every activation of every cell is compared.

    Measured on MI355X: V RMS 0.137 mV over the 2000 ms; the first
    activation of every cell (the S1 wave) within the 0.1 ms bar; of the
    2.1 million activations of cells that activate equally often in both
    precisions, 6.4% differ by more than 0.1 ms (99.9th percentile 0.57
    ms, largest 2.2 ms): the re-entrant wave arrives in tissue that is
    still recovering, where a small difference in the gates' recovery moves
    the upstroke by much more; 145 cells of 1,048,576, at the edge of the
    conduction block, activate once more or less in one precision. The
    bars below hold the single-precision bars where they apply (RMS, the
    first activation) and guard the rest against regressions."""
    t0 = time.time()
    rms, acts, seen = _config2_run(440.0)
    c = _compare_activations(acts, seen)
    n = seen[64].size
    print('config2 1024^2 S1 + S2 at 440 ms, 2000 ms (%.0f s): fp32 vs fp64 '
          'V RMS %.3g mV; first activation max|dt| %.3g ms; over the %d '
          'activations of cells that activate equally often: max|dt| %.3g '
          'ms, 99.9th percentile %.3g ms, %d beyond 0.1 ms; cells with '
          'different activation counts %d of %d; activations after the first'
          ' (fp64) %d; most activations of one cell %d'
          % (time.time() - t0, rms, c['first'], c['n'], c['dact'], c['p999'],
             c['over'], c['differ'], n, c['later'], c['most']), flush=True)
    assert c['later'] > 0
    assert rms <= 0.5
    assert c['first'] <= 0.1
    assert c['differ'] <= 1e-3 * n
    assert c['p999'] <= 1.0 and c['dact'] <= 5.0
    assert c['over'] <= 0.1 * c['n']


@pytest.mark.parametrize('driver', ['block', 'grid'])
def test_conductance_field_fp32(driver):
    """Single-precision conductance fields (diff_hetero) against the
    double-precision oracle, 30 x 24 sheet, corner stimulus, 30 ms."""
    rng = np.random.default_rng(8)
    nx, ny = 30, 24
    gx = rng.uniform(5, 12, size=(ny, nx - 1))
    gy = rng.uniform(5, 12, size=(ny - 1, nx))
    sc = Scenario('lr1991', (nx, ny), rl=True, dt=0.01, paced=(4, 4, 0, 0),
                  conductance_field=(gx, gy))
    s = sc.product(myokit.SINGLE_PRECISION)
    s.set_tuning(driver=driver)
    b = s.run_block(30.0, ['membrane.V'], log_interval=0.1)
    assert s.engine_driver() == driver
    r = sc.oracle().run(30.0, log_vars=['membrane.V'], log_interval=0.1)
    V = b['membrane.V'].reshape(len(b['time']), -1)
    rms, dact = _fp32_report('hetero ' + driver, b['time'], V,
                             r['membrane.V'])
    assert rms <= 0.5 and dact <= 0.1


def test_native_maths_fp32_sheet():
    """native_maths=True in single precision (hardware exp/log, the
    reference's native_* functions) on a 256 x 256 RL sheet (persistent
    sheet driver), S1 on the left 4 columns, 30 ms, against the
    double-precision oracle."""
    sc = Scenario('lr1991', (256, 256), rl=True, dt=0.01,
                  paced=(4, 256, 0, 0))
    s = myokit.SimulationOpenCL(
        myokit.load_model('example'), sc.protocol(), ncells=(256, 256),
        rl=True, precision=myokit.SINGLE_PRECISION, native_maths=True)
    s.set_step_size(0.01)
    s.set_conductance(9.5, 9.5)
    s.set_paced_cells(4, 256, 0, 0)
    b = s.run_block(30.0, ['membrane.V'], log_interval=0.1)
    assert s.engine_driver() == 'sheet'
    r = sc.oracle().run(30.0, log_vars=['membrane.V'], log_interval=0.1)
    V = b['membrane.V'].reshape(len(b['time']), -1)
    rms, dact = _fp32_report('native maths fp32 256^2', b['time'], V,
                             r['membrane.V'])
    assert rms <= 0.5 and dact <= 0.1
