"""
One GPU check per benchmark configuration, each printing its accuracy:
double precision max|dV| against the oracle, single precision V RMS error
and activation-time error against the double-precision oracle (bars:
1e-6 mV; 0.5 mV RMS and 0.1 ms). Where the CPU oracle would need minutes,
the test runs a stated reduction of the configuration (fewer cells or a
shorter window); the reduction is named in each docstring.
"""
import numpy as np
import pytest

import myokit_amd as myokit
from helpers import Scenario, activation_times, run_both, run_product

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-9


def _fp32_report(name, P, O, t_act=True):
    dV = P['V'] - O['V']
    rms = float(np.sqrt(np.mean(dV ** 2)))
    line = '%s fp32: V RMS %.3g mV, max|dV| %.3g mV' % (
        name, rms, float(np.max(np.abs(dV))))
    dact = 0.0
    if t_act:
        a = activation_times(P['time'], P['V'])
        b = activation_times(O['time'], O['V'])
        both = np.isfinite(a) & np.isfinite(b)
        assert np.array_equal(np.isfinite(a), np.isfinite(b))
        dact = float(np.max(np.abs(a[both] - b[both]))) if both.any() else 0.0
        line += ', activation max|dt| %.3g ms over %d cells' % (
            dact, int(both.sum()))
    print(line, flush=True)
    return rms, dact


def _fp64_report(name, P, O):
    d = float(np.max(np.abs(P['V'] - O['V'])))
    print('%s fp64: max|dV| %.3g mV (rows %d, cells %d)'
          % (name, d, P['V'].shape[0], P['V'].shape[1]), flush=True)
    return d


def test_config1_cable_full_run():
    """Config 1 in full: 200-cell LR1991 cable, FE, dt 0.005 ms, 1000 ms,
    gx 9.5, cells 0-4 paced at 1 Hz (2 ms), V logged every 1 ms."""
    sc = Scenario('lr1991', 200, dt=0.005, gx=9.5,
                  events=[(1, 0, 2, 1000, 0)], paced=(5,))
    P, O = run_both(sc, 1000.0, myokit.DOUBLE_PRECISION, log_interval=1.0)
    assert P['V'].shape == (1000, 200)
    assert _fp64_report('config1', P, O) <= FP64_TOL
    P32 = run_product(sc, 1000.0, myokit.SINGLE_PRECISION, log_interval=1.0)
    rms, dact = _fp32_report('config1', P32, O)
    assert rms <= 0.5 and dact <= 0.1


@pytest.mark.parametrize('precision', [64, 32])
def test_config2_sheet_reduced(precision):
    """Config 2 reduced to a 256 x 256 sheet and the first 30 ms (S1 wave
    crossing the sheet); LR1991, RL, dt 0.01 ms, gx = gy = 9.5, S1 on the
    left 4 columns. V logged every 0.1 ms for activation times."""
    sc = Scenario('lr1991', (256, 256), rl=True, dt=0.01, gx=9.5, gy=9.5,
                  events=[(1, 0, 2, 1000, 0)], paced=(4, 256, 0, 0))
    P, O = run_both(sc, 30.0, precision, log_interval=0.1)
    assert P['sim'].engine_driver() == 'sheet'
    if precision == 64:
        assert _fp64_report('config2 256^2', P, O) <= FP64_TOL
    else:
        rms, dact = _fp32_report('config2 256^2', P, O)
        assert rms <= 0.5 and dact <= 0.1


@pytest.mark.parametrize('precision', [64, 32])
def test_config2_full_size_window(precision):
    """Config 2 at full size (1024 x 1024), first 2 ms (the S1 stimulus),
    V of every cell logged every 0.5 ms."""
    sc = Scenario('lr1991', (1024, 1024), rl=True, dt=0.01, gx=9.5, gy=9.5,
                  events=[(1, 0, 2, 1000, 0)], paced=(4, 1024, 0, 0))
    s = sc.product(precision)
    b = s.run_block(2.0, ['membrane.V'], log_interval=0.5)
    # in double precision 1024 x 1024 tiles are not co-resident: grid driver
    assert s.engine_driver() == ('sheet' if precision == 32 else 'grid')
    P = {'time': b['time'], 'V': b['membrane.V'].reshape(len(b['time']), -1)}
    r = sc.oracle().run(2.0, log_vars=['membrane.V'], log_interval=0.5)
    O = {'time': r['time'], 'V': r['membrane.V']}
    assert P['V'].shape == O['V'].shape == (4, 1024 * 1024)
    if precision == 64:
        assert _fp64_report('config2 1024^2', P, O) <= FP64_TOL
    else:
        rms, _ = _fp32_report('config2 1024^2', P, O, t_act=False)
        assert rms <= 0.5


def _scales(rng, n):
    return np.exp(rng.uniform(np.log(0.5), np.log(2.0), n))


@pytest.mark.parametrize('precision', [64, 32])
def test_config3_uncoupled_batch_reduced(precision):
    """Config 3(i) reduced to 2048 cells and 400 ms: uncoupled LR1991
    cells (diffusion off, all paced), gNa, gCa (gsi) and gK scaled by
    factors drawn log-uniform in [0.5, 2] from a seed, FE, dt 0.005 ms,
    V logged every 0.5 ms."""
    m = myokit.load_model('example')
    rng = np.random.default_rng(3)
    n = 2048
    fields = {}
    for name in ('ina.gNa', 'ica.gCa', 'ik.IK.gK'):
        fields[name] = m.get(name).rhs().eval() * _scales(rng, n)
    sc = Scenario('lr1991', n, diffusion=False, dt=0.005,
                  events=[(1, 0, 2, 1000, 0)], fields=fields)
    P, O = run_both(sc, 400.0, precision, log_interval=0.5)
    assert P['sim'].engine_driver() == 'cells'
    if precision == 64:
        assert _fp64_report('config3 cells', P, O) <= FP64_TOL
    else:
        rms, dact = _fp32_report('config3 cells', P, O)
        assert rms <= 0.5 and dact <= 0.1


@pytest.mark.parametrize('precision', [64, 32])
def test_config3_uneven_cables_reduced(precision):
    """Config 3(ii) reduced to 12 cables with lengths drawn uniformly in
    [50, 400] (seeded) and 60 ms: chains built with set_connections
    (g = 9.5), the first 5 cells of each cable paced, FE, dt 0.005 ms, all
    cables in one launch (the chain driver: one-wave tiles laid end to
    end)."""
    rng = np.random.default_rng(11)
    lengths = rng.integers(50, 401, 12)
    conns, paced = [], []
    off = 0
    for L in lengths:
        conns += [(off + i, off + i + 1, 9.5) for i in range(L - 1)]
        paced += list(range(off, off + 5))
        off += L
    sc = Scenario('lr1991', int(off), dt=0.005, connections=conns,
                  paced_list=paced, events=[(1, 0, 2, 1000, 0)])
    P, O = run_both(sc, 60.0, precision, log_interval=0.5)
    assert P['sim'].engine_driver() == 'chain'
    if precision == 64:
        assert _fp64_report('config3 cables', P, O) <= FP64_TOL
    else:
        rms, dact = _fp32_report('config3 cables', P, O)
        assert rms <= 0.5 and dact <= 0.1


def _smooth_field(rng, n, lo=0.7, hi=1.3):
    """Sum of 8 random sinusoids rescaled to [lo, hi]."""
    x = np.arange(n) / n
    s = np.zeros(n)
    for _ in range(8):
        s += rng.uniform(0.2, 1.0) * np.sin(
            2 * np.pi * rng.uniform(0.5, 6.0) * x + rng.uniform(0, 2 * np.pi))
    s = (s - s.min()) / (s.max() - s.min())
    return lo + (hi - lo) * s


def _config4_sim(seed, gx):
    m = myokit.load_model('example')
    rng = np.random.default_rng(seed)
    s = myokit.SimulationOpenCL(m, myokit.pacing.blocktrain(1000, 2),
                                ncells=1000, precision=64)
    s.set_step_size(0.005)
    s.set_conductance(gx)
    s.set_paced_cells(5)
    s.set_field('ina.gNa', 16.0 * _smooth_field(rng, 1000))
    s.set_field('ica.gCa', 0.09 * _smooth_field(rng, 1000))
    return s


def test_config4_adjoint_full():
    """Config 4 in full: 1000-cell LR1991 cable, FP64, FE, dt 0.005 ms,
    500 ms (100,000 steps); gNa and gCa scale fields smooth in [0.7, 1.3];
    target from a hidden seeded parameter set (other fields, gx = 11);
    loss over 20 evenly spaced probes; checkpoints every 200 steps.
    Gradients against central differences: gx with h = 1e-6 (relative), two
    field entries with h = 1e-4 (relative; the effect of one cell's
    conductance on the loss is small, and a larger step keeps it above the
    rounding noise of a 2e6-term loss)."""
    from myokit_amd.adjoint import TissueAdjoint
    probes = np.linspace(25, 975, 20).astype(int)
    dur = 500.0
    hidden = _config4_sim(101, 11.0)
    target = TissueAdjoint(hidden, probes).trace(dur)
    s = _config4_sim(7, 9.5)
    adj = TissueAdjoint(s, probes, checkpoint_every=200)
    assert adj.n_steps(dur) == 100000 == target.shape[0]
    loss, g, dgx, _ = adj.loss_and_gradient(dur, target)
    print('config4: loss %.10g dL/dgx %.10g phases %s'
          % (loss, dgx, adj.timing()), flush=True)

    def L(gx=9.5, name=None, c=None, v=None):
        s.set_conductance(gx)
        if name is not None:
            f = np.array(s._fields[s._model.get(name)])
            f0 = f[c]
            f[c] = v
            s.set_field(name, f)
        r = adj.loss_and_gradient(dur, target)[0]
        s.set_conductance(9.5)
        if name is not None:
            f[c] = f0
            s.set_field(name, f)
        return r
    h = 9.5e-6
    d = (L(gx=9.5 + h) - L(gx=9.5 - h)) / (2 * h)
    print('config4: dL/dgx adjoint %.10g fd %.10g' % (dgx, d), flush=True)
    assert abs(dgx - d) <= 1e-5 * abs(d)
    for name, c in (('ina.gNa', int(probes[3])), ('ica.gCa', 500)):
        v0 = float(s._fields[s._model.get(name)][c])
        h = 1e-4 * v0
        d = (L(name=name, c=c, v=v0 + h) - L(name=name, c=c, v=v0 - h)) \
            / (2 * h)
        print('config4: dL/d%s[%d] adjoint %.10g fd %.10g'
              % (name, c, g[name][c], d), flush=True)
        assert abs(g[name][c] - d) <= 1e-5 * abs(d)


def _geometric_graph(rng, n, degree=6.0):
    """Random geometric graph in the unit square with mean degree
    ``degree``; conductances uniform in [5, 12]."""
    from scipy.spatial import cKDTree
    pts = rng.uniform(0, 1, (n, 2))
    r = np.sqrt(degree / (np.pi * n))
    pairs = np.array(sorted(cKDTree(pts).query_pairs(r)))
    g = rng.uniform(5, 12, len(pairs))
    return pts, [(int(i), int(j), float(c)) for (i, j), c in zip(pairs, g)]


def test_config5_batched_graph_adjoint_reduced():
    """Config 5 reduced to 8,192 cells, 2 parameter sets and 20 ms: a
    seeded random geometric graph (mean degree 6, g uniform in [5, 12]),
    RL, dt 0.01 ms, fields gNa, gCa, gK scaled log-uniform in [0.5, 2] per
    set. The sets run as one batch: the disjoint union of two copies of the
    graph, one per set. Single-precision state with double-precision loss
    and gradient reductions, against the same run in double precision:
    with the default double-precision VJP arithmetic (bounded) and with a
    single-precision VJP (error reported)."""
    from myokit_amd.adjoint import TissueAdjoint
    rng = np.random.default_rng(21)
    n = 4096
    pts, conns = _geometric_graph(rng, n)
    sets = 2
    allc = [(i + k * n, j + k * n, g) for k in range(sets)
            for i, j, g in conns]
    paced = [k * n + i for k in range(sets) for i in np.nonzero(
        pts[:, 0] < 0.05)[0]]
    m = myokit.load_model('example')
    fields = {name: m.get(name).rhs().eval() * _scales(rng, sets * n)
              for name in ('ina.gNa', 'ica.gCa', 'ik.IK.gK')}
    probes = [k * n + i for k in range(sets) for i in range(0, n, 512)]
    out = {}
    for prec in (64, 32, 'f64-vjp'):
        s = myokit.SimulationOpenCL(m, myokit.pacing.blocktrain(1000, 2),
                                    ncells=sets * n, rl=True,
                                    precision=64 if prec == 64 else 32)
        s.set_step_size(0.01)
        s.set_connections(allc)
        s.set_paced_cell_list(paced)
        for name, v in fields.items():
            s.set_field(name, v)
        adj = TissueAdjoint(s, probes, checkpoint_every=500,
                            vjp_precision=64 if prec == 'f64-vjp' else None)
        target = np.full((adj.n_steps(20.0), len(probes)), -60.0)
        out[prec] = adj.loss_and_gradient(20.0, target)
    l64, g64, s64, _ = out[64]
    for kind, bound in ((32, 1e-3), ('f64-vjp', 1e-3)):
        l32, g32, s32, _ = out[kind]
        print('config5 %s: loss %.8g vs fp64 %.8g; dL/dscale %.8g vs %.8g'
              ' (relative error %.3g)' % (kind, l32, l64, s32, s64,
                                          abs(s32 - s64) / abs(s64)),
              flush=True)
        assert abs(l32 - l64) <= 1e-3 * abs(l64)
        errs = [abs(s32 - s64) / abs(s64)]
        for name in g64:
            e = np.linalg.norm(g32[name] - g64[name]) / \
                np.linalg.norm(g64[name])
            errs.append(e)
            print('config5 %s: %s gradient relative error %.3g'
                  % (kind, name, e), flush=True)
        assert max(errs) <= bound
