"""
GPU tests of the discrete adjoint (myokit_amd.adjoint): gradients against
central finite differences of the same discrete loss (h = 1e-6, relative),
against a closed form for a linear model, single against double precision,
and the torch.autograd bridge.
"""
import numpy as np
import pytest

import myokit_amd as myokit
from myokit_amd.adjoint import TissueAdjoint

pytestmark = pytest.mark.gpu


def cable(n=30, precision=64, rl=False, dt=0.005, seed=1):
    m = myokit.load_model('example')
    s = myokit.SimulationOpenCL(m, myokit.pacing.blocktrain(1000, 2),
                                ncells=n, precision=precision, rl=rl)
    s.set_step_size(dt)
    s.set_conductance(9.5)
    rng = np.random.default_rng(seed)
    s.set_field('ina.gNa', 16 * rng.uniform(0.8, 1.2, n))
    s.set_field('ica.gCa', 0.09 * rng.uniform(0.8, 1.2, n))
    return s


def fd(f, x, h):
    return (f(x + h) - f(x - h)) / (2 * h)


def rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def test_adjoint_cable_fd_fp64():
    s = cable()
    probes = [3, 10, 25]
    adj = TissueAdjoint(s, probes, checkpoint_every=64)
    dur = 4.0
    # Target: the trace of a hidden parameter set
    hidden = cable(seed=2)
    hidden.set_conductance(11.0)
    target = TissueAdjoint(hidden, probes).trace(dur)
    assert target.shape == (adj.n_steps(dur), 3)
    loss, g, dgx, dgy = adj.loss_and_gradient(dur, target)
    # The loss is that of the forward trace
    tr = adj.trace(dur)
    assert loss == pytest.approx(np.sum((tr - target) ** 2), rel=1e-12)

    def loss_gx(gx):
        s.set_conductance(gx)
        r = adj.loss_and_gradient(dur, target)[0]
        s.set_conductance(9.5)
        return r
    d = fd(loss_gx, 9.5, 9.5e-6)
    print('dL/dgx adjoint %.10g fd %.10g' % (dgx, d), flush=True)
    assert rel(dgx, d) < 1e-5

    base = {k: np.array(v) for k, v in
            ((var.qname(), vals) for var, vals in s._fields.items())}
    for name in ('ina.gNa', 'ica.gCa'):
        for c in (2, 11, 26):
            def loss_f(v):
                x = base[name].copy()
                x[c] = v
                s.set_field(name, x)
                r = adj.loss_and_gradient(dur, target)[0]
                s.set_field(name, base[name])
                return r
            v0 = base[name][c]
            # Relative step 1e-4: truncation error ~1e-8 relative, while the
            # rounding noise of the loss (~1e-16 L / h) stays far below the
            # smallest gradients checked
            h = v0 * 1e-4
            d = fd(loss_f, v0, h)
            print('dL/d%s[%d] adjoint %.10g fd %.10g (L %.6g)'
                  % (name, c, g[name][c], d, loss), flush=True)
            assert rel(g[name][c], d) < 1e-5 or \
                abs(g[name][c] - d) < 1e-12 * loss / h


def test_adjoint_sheet_rl_fd_fp64():
    m = myokit.load_model('example')
    s = myokit.SimulationOpenCL(m, myokit.pacing.blocktrain(1000, 2),
                                ncells=(12, 10), precision=64, rl=True)
    s.set_step_size(0.01)
    s.set_conductance(9.0, 7.0)
    s.set_paced_cells(2, 10, 0, 0)
    rng = np.random.default_rng(4)
    s.set_field('ina.gNa', 16 * rng.uniform(0.8, 1.2, (10, 12)))
    probes = [(5, 5), (10, 2), (11, 9)]
    adj = TissueAdjoint(s, probes, checkpoint_every=50)
    dur = 6.0
    target = np.full((adj.n_steps(dur), 3), -20.0)
    loss, g, dgx, dgy = adj.loss_and_gradient(dur, target)

    def lg(gx, gy):
        s.set_conductance(gx, gy)
        r = adj.loss_and_gradient(dur, target)[0]
        s.set_conductance(9.0, 7.0)
        return r
    dx = fd(lambda v: lg(v, 7.0), 9.0, 9e-6)
    dy = fd(lambda v: lg(9.0, v), 7.0, 7e-6)
    print('sheet dgx %.10g fd %.10g; dgy %.10g fd %.10g' % (dgx, dx, dgy, dy),
          flush=True)
    assert rel(dgx, dx) < 1e-5 and rel(dgy, dy) < 1e-5
    base = np.array(s._fields[s._model.get('ina.gNa')]).reshape(10, 12)
    for (x, y) in [(5, 5), (3, 4)]:
        def lf(v):
            b = base.copy()
            b[y, x] = v
            s.set_field('ina.gNa', b)
            r = adj.loss_and_gradient(dur, target)[0]
            s.set_field('ina.gNa', base)
            return r
        d = fd(lf, base[y, x], base[y, x] * 1e-4)
        print('sheet dgNa[%d,%d] %.10g fd %.10g' % (y, x, g['ina.gNa'][y, x],
                                                   d), flush=True)
        assert rel(g['ina.gNa'][y, x], d) < 1e-5


def test_adjoint_graph_scale_fd_fp64():
    rng = np.random.default_rng(9)
    n = 40
    conns = [(i, i + 1, float(rng.uniform(5, 12))) for i in range(n - 1)]
    conns += [(0, 20, 6.0), (5, 33, 8.0)]
    m = myokit.load_model('example')
    s = myokit.SimulationOpenCL(m, myokit.pacing.blocktrain(1000, 2),
                                ncells=n, precision=64)
    s.set_connections(conns)
    s.set_paced_cell_list([0, 1])
    adj = TissueAdjoint(s, [10, 30], checkpoint_every=100)
    dur = 5.0
    target = np.full((adj.n_steps(dur), 2), -50.0)
    loss, g, ds, _ = adj.loss_and_gradient(dur, target)

    def ls(f):
        s.set_connections([(i, j, c * f) for i, j, c in conns])
        r = adj.loss_and_gradient(dur, target)[0]
        s.set_connections(conns)
        return r
    d = fd(ls, 1.0, 1e-6)
    print('graph dL/dscale %.10g fd %.10g' % (ds, d), flush=True)
    assert rel(ds, d) < 1e-5


def test_adjoint_linear_model_closed_form():
    # dot(x) = -k x + u with per-cell fields k and u, forward Euler:
    # x_{n+1} = (1 - k dt) x_n + dt u; dx/dk and dx/du by recursion
    src = '''[[model]]
c.x = 1.5
[c]
dot(x) = -k * x + u
k = 0.5
u = 0.2
t = 0 bind time
'''
    m = myokit.parse_model(src)
    n = 4
    s = myokit.SimulationOpenCL(m, None, ncells=n, diffusion=False,
                                precision=64)
    dt = 0.01
    s.set_step_size(dt)
    k = np.array([0.5, 1.0, 2.0, 0.1])
    u = np.array([0.2, -0.3, 1.0, 0.0])
    s.set_field('c.k', k)
    s.set_field('c.u', u)
    adj = TissueAdjoint(s, [0, 1, 2, 3], checkpoint_every=7)
    N = adj.n_steps(1.0)
    assert N == 100
    rng = np.random.default_rng(0)
    target = rng.uniform(0, 1, (N, n))
    loss, g, _, _ = adj.loss_and_gradient(1.0, target)
    x = np.full(n, 1.5)
    dk = np.zeros(n)
    du = np.zeros(n)
    L = 0.0
    gk = np.zeros(n)
    gu = np.zeros(n)
    for i in range(N):
        dk, du = (1 - k * dt) * dk - dt * x, (1 - k * dt) * du + dt
        x = (1 - k * dt) * x + dt * u
        r = x - target[i]
        L += np.sum(r * r)
        gk += 2 * r * dk
        gu += 2 * r * du
    assert loss == pytest.approx(L, rel=1e-12)
    np.testing.assert_allclose(g['c.k'], gk, rtol=1e-10)
    np.testing.assert_allclose(g['c.u'], gu, rtol=1e-10)


@pytest.mark.parametrize('vjp', [64, 32])
def test_adjoint_single_precision_against_double(vjp):
    # Single-precision state; the VJP in single (default) or double
    dur = 3.0
    probes = [5, 20]
    out = {}
    for prec in (32, 64):
        s = cable(n=30, precision=prec)
        adj = TissueAdjoint(s, probes, checkpoint_every=100,
                            vjp_precision=vjp)
        target = np.full((adj.n_steps(dur), 2), -30.0)
        out[prec] = adj.loss_and_gradient(dur, target)
    l32, g32, x32, _ = out[32]
    l64, g64, x64, _ = out[64]
    print('vjp %d: fp32 loss %.8g fp64 %.8g; dgx %.8g / %.8g'
          % (vjp, l32, l64, x32, x64), flush=True)
    assert rel(l32, l64) < 1e-4
    tol = 2e-3
    assert rel(x32, x64) < tol
    for name in g64:
        err = np.max(np.abs(g32[name] - g64[name])) / np.max(np.abs(g64[name]))
        print('  %s max error %.3g' % (name, err), flush=True)
        assert err < tol, name


def test_torch_autograd_bridge():
    torch = pytest.importorskip('torch')
    from myokit_amd.adjoint import TissueLoss
    s = cable(n=20)
    adj = TissueAdjoint(s, [4, 15], checkpoint_every=50)
    dur = 2.0
    target = torch.full((adj.n_steps(dur), 2), -40.0, dtype=torch.float64)
    gna = torch.tensor(16 * np.linspace(0.9, 1.1, 20), requires_grad=True)
    gca = torch.tensor(0.09 * np.ones(20), requires_grad=True)
    gx = torch.tensor(9.5, dtype=torch.float64, requires_grad=True)
    gy = torch.tensor(5.0, dtype=torch.float64, requires_grad=True)
    loss = TissueLoss.apply(adj, dur, target, gx, gy, gna, gca)
    loss.backward()
    l2, g, dgx, _ = adj.loss_and_gradient(dur, target.numpy())
    assert float(loss.detach()) == pytest.approx(l2, rel=1e-14)
    np.testing.assert_allclose(gna.grad.numpy(), g['ina.gNa'], rtol=1e-12)
    np.testing.assert_allclose(gca.grad.numpy(), g['ica.gCa'], rtol=1e-12)
    assert float(gx.grad) == pytest.approx(dgx, rel=1e-12)


@pytest.mark.parametrize('K', [1000, 70])
def test_adjoint_sheet_driver_matches_grid_fp64(K):
    # 128 x 64 cells: large enough for the persistent sheet driver, whose
    # forward and recomputation passes store the trajectory from registers.
    # In double precision the sheet is bit-exact with the grid driver, so
    # the loss and the field gradients are identical; the reverse pass runs
    # on the persistent reverse sheet (one workgroup per CU in double
    # precision), whose conductance partial sums are grouped by tile, so
    # dL/dgx and dL/dgy agree to 1e-12. K = 70 exercises recomputation.
    out = []
    for force_grid in (False, True):
        m = myokit.load_model('example')
        s = myokit.SimulationOpenCL(m, myokit.pacing.blocktrain(1000, 2),
                                    ncells=(128, 64), precision=64, rl=True)
        s.set_step_size(0.01)
        s.set_conductance(9.5, 6.0)
        s.set_paced_cells(3, 64, 0, 0)
        s.set_field('ina.gNa', np.linspace(14, 18, 128 * 64).reshape(64, 128))
        if force_grid:
            s.set_engine_option('driver', 2)
        adj = TissueAdjoint(s, [(10, 10), (60, 30), (127, 63)],
                            checkpoint_every=K)
        dur = 3.0
        target = np.full((adj.n_steps(dur), 3), -40.0)
        out.append(adj.loss_and_gradient(dur, target))
        print('driver', s.engine_driver(), 'reverse', adj.reverse_kernel(),
              'loss', out[-1][0], 'dgx', out[-1][2], 'dgy', out[-1][3],
              flush=True)
        assert s.engine_driver() == ('grid' if force_grid else 'sheet')
        # with recomputation (K = 70) the 8,192-cell grid path overlaps it
        # with the reverse steps on a second stream
        assert adj.reverse_kernel()[0] == (
            ('step' if K >= 300 else 'step-overlap') if force_grid
            else 'sheet')
    (l1, g1, x1, y1), (l2, g2, x2, y2) = out
    assert l1 == l2
    np.testing.assert_array_equal(g1['ina.gNa'], g2['ina.gNa'])
    assert x1 == pytest.approx(x2, rel=1e-12)
    assert y1 == pytest.approx(y2, rel=1e-12)


@pytest.mark.parametrize('K', [10000, 37])
def test_adjoint_block_layout_matches_grid_fp64(K):
    # Cables and graph components run the adjoint on the block layout (one
    # workgroup per domain, every reverse step of a segment per launch, abar
    # carried between segments). Against the one-step-per-launch grid path:
    # the same arithmetic per cell, only the conductance partial sums are
    # added in another order.
    def make(graph, force_grid):
        m = myokit.load_model('example')
        n = 60
        s = myokit.SimulationOpenCL(m, myokit.pacing.blocktrain(1000, 2),
                                    ncells=n, precision=64)
        s.set_step_size(0.005)
        rng = np.random.default_rng(5)
        if graph:
            conns = [(i, i + 1, float(rng.uniform(5, 12)))
                     for i in range(n - 1) if i != 29]
            conns += [(3, 17, 7.0), (40, 55, 6.0)]
            s.set_connections(conns)
            s.set_paced_cell_list([0, 1, 30, 31])
        else:
            s.set_conductance(9.5)
        s.set_field('ina.gNa', 16 * rng.uniform(0.8, 1.2, n))
        if force_grid:
            s.set_engine_option('driver', 2)
        return s
    for graph in (False, True):
        res = []
        for force_grid in (False, True):
            s = make(graph, force_grid)
            adj = TissueAdjoint(s, [5, 33, 58], checkpoint_every=K)
            target = np.full((adj.n_steps(4.0), 3), -50.0)
            res.append(adj.loss_and_gradient(4.0, target))
            assert s.engine_driver() == ('grid' if force_grid else 'block')
        (l1, g1, x1, _), (l2, g2, x2, _) = res
        print('graph' if graph else 'cable', 'K', K, 'loss', l1, l2,
              'dg', x1, x2, flush=True)
        assert l1 == pytest.approx(l2, rel=1e-13)
        assert x1 == pytest.approx(x2, rel=1e-11)
        np.testing.assert_allclose(g1['ina.gNa'], g2['ina.gNa'], rtol=1e-11,
                                   atol=1e-12 * np.max(np.abs(g2['ina.gNa'])))


def _sheet_adjoint(shape, K, precision, sheet, vjp=None, dur=3.0):
    nx, ny = shape
    m = myokit.load_model('example')
    s = myokit.SimulationOpenCL(m, myokit.pacing.blocktrain(1000, 2),
                                ncells=shape, precision=precision, rl=True)
    s.set_step_size(0.01)
    s.set_conductance(9.5, 6.0)
    s.set_paced_cells(3, ny, 0, 0)
    s.set_field('ina.gNa', np.linspace(12, 20, nx * ny).reshape(ny, nx))
    s.set_engine_option('adj_sheet', sheet)
    kw = {} if vjp is None else {'vjp_precision': vjp}
    adj = TissueAdjoint(s, [(10, 10), (nx // 2, ny // 2), (nx - 1, ny - 1)],
                        checkpoint_every=K, **kw)
    target = np.full((adj.n_steps(dur), 3), -40.0)
    loss, g, gx, gy = adj.loss_and_gradient(dur, target)
    assert s.engine_driver() == 'sheet'
    kind, launches = adj.reverse_kernel()
    print('precision', precision, 'adj_sheet', sheet, 'reverse', kind,
          launches, 'loss', loss, 'dgx', gx, 'dgy', gy, flush=True)
    return loss, np.asarray(g['ina.gNa']), gx, gy, kind, launches


SHEET_CASES = [((128, 96), 10000), ((100, 70), 37), ((100, 70), 299)]


@pytest.mark.parametrize('shape,K', SHEET_CASES)
def test_adjoint_persistent_sheet_matches_step_kernel_fp64(shape, K):
    # Sheets run the reverse sweep in one persistent cooperative launch per
    # segment (mk_adj_sheet: mu in LDS, abar exchanged between tiles through
    # edge buffers and step counters). Per cell it does the arithmetic of
    # the one-launch-per-step kernel in the same order, so in double
    # precision the loss and the field gradients are identical bit for bit:
    # any stale or misplaced abar from the exchange would show. The
    # conductance gradients differ only in how partial sums are grouped.
    # (100, 70) has partial tiles in x and y; K = 37 gives several segments,
    # K = 299 a one-step segment (300 steps).
    nsteps = 300
    nseg = -(-nsteps // min(K, nsteps))
    l1, g1, x1, y1, k1, n1 = _sheet_adjoint(shape, K, 64, 1)
    l2, g2, x2, y2, k2, n2 = _sheet_adjoint(shape, K, 64, 0)
    assert (k1, n1) == ('sheet', nseg)
    assert (k2, n2) == ('step', nsteps + 1)
    assert l1 == l2
    np.testing.assert_array_equal(g1, g2)
    assert x1 == pytest.approx(x2, rel=1e-12)
    assert y1 == pytest.approx(y2, rel=1e-12)


@pytest.mark.parametrize('shape,K', SHEET_CASES)
def test_adjoint_persistent_sheet_fp32(shape, K):
    # In single precision too the persistent and the one-launch-per-step
    # reverse kernels run the same arithmetic, so the field gradients are
    # identical. (A read of a transcendental's result by the inline-assembly
    # multiply without its wait state once made them differ by 1e-8, which
    # the upstroke amplified to ~1e-6 of the largest entry.) Both are also
    # held to the gradient with a double-precision cell VJP (same
    # single-precision forward): within 1e-5 of its largest entry, dL/dgx
    # and dL/dgy within 1e-5 relative. The loss is the forward's.
    nsteps = 300
    nseg = -(-nsteps // min(K, nsteps))
    l1, g1, x1, y1, k1, n1 = _sheet_adjoint(shape, K, 32, 1)
    l2, g2, x2, y2, k2, n2 = _sheet_adjoint(shape, K, 32, 0)
    l3, g3, x3, y3, k3, n3 = _sheet_adjoint(shape, K, 32, 0, vjp=64)
    assert (k1, n1) == ('sheet', nseg)
    assert k2 == 'step' and k3 == 'step'
    assert l1 == l2 == l3
    np.testing.assert_array_equal(g1, g2)
    scale = np.max(np.abs(g3))
    e1 = np.max(np.abs(g1 - g3)) / scale
    e2 = np.max(np.abs(g2 - g3)) / scale
    print('fp32 field gradient vs f64 VJP: sheet %.3g, step %.3g' % (e1, e2),
          flush=True)
    assert e1 <= 1e-5 and e2 <= 1e-5
    for a in (x1, x2):
        assert a == pytest.approx(x3, rel=1e-5)
    for a in (y1, y2):
        assert a == pytest.approx(y3, rel=1e-5)


def test_adjoint_graph_cell_order_fd_fp64():
    # A random geometric graph is held on the device in another cell order;
    # the field gradients are gathered back into the caller's numbering on
    # the device. Central differences of the loss in gNa of single cells,
    # numbered as the caller numbers them, pin that mapping; then again
    # after set_connections changes the graph (and the device order).
    from scipy.spatial import cKDTree
    n = 300
    rng = np.random.default_rng(17)
    pts = rng.uniform(0, 1, (n, 2))
    m = myokit.load_model('example')
    s = myokit.SimulationOpenCL(m, myokit.pacing.blocktrain(1000, 2),
                                ncells=n, precision=64)
    s.set_step_size(0.005)

    def graph(r):
        pairs = cKDTree(pts).query_pairs(r, output_type='ndarray')
        return [(int(i), int(j), float(g)) for (i, j), g in
                zip(pairs, rng.uniform(5, 12, len(pairs)))]
    paced = [int(c) for c in np.nonzero(pts[:, 0] < 0.15)[0]]
    s.set_paced_cell_list(paced)
    gna = 16 * rng.uniform(0.8, 1.2, n)
    s.set_field('ina.gNa', gna)
    probes = [int(np.argmin(np.abs(pts[:, 0] - 0.3))),
              int(np.argmin(np.abs(pts[:, 0] - 0.5)))]
    dur = 4.0
    for r in (0.09, 0.11):
        s.set_connections(graph(r))
        perm, iperm = s._cell_order()
        assert perm is not None and not np.array_equal(perm, np.arange(n))
        adj = TissueAdjoint(s, probes, checkpoint_every=300)
        target = np.full((adj.n_steps(dur), 2), -45.0)
        loss, g, _, _ = adj.loss_and_gradient(dur, target)
        g = np.array(g['ina.gNa'])
        # probe cells and the largest gradient elsewhere
        others = np.argsort(-np.abs(g))
        cells = probes + [int(c) for c in others if c not in probes][:1]
        for c in cells:
            def lc(v, c=c):
                f = gna.copy()
                f[c] = v
                s.set_field('ina.gNa', f)
                return adj.loss_and_gradient(dur, target)[0]
            d = fd(lc, gna[c], 1e-6 * gna[c])
            s.set_field('ina.gNa', gna)
            print('graph r %g cell %d (device %d) dL/dgNa %.10g fd %.10g'
                  % (r, c, iperm[c], g[c], d), flush=True)
            assert rel(g[c], d) < 1e-5


def test_adjoint_graph_copies_equal_plain_layout_fp64():
    # A batch of four identical copies of a 2,500-cell graph: held
    # interleaved with one copy's connection rows, the loss and every field
    # gradient must equal, bit for bit, those of the same simulation held
    # in the plain reverse Cuthill-McKee order with a row per cell (the
    # same operations in every cell, only where the rows come from
    # differs)
    from scipy.spatial import cKDTree
    from myokit_amd import simulation
    m, K = 2500, 4
    rng = np.random.default_rng(23)
    pts = rng.uniform(0, 1, (m, 2))
    pairs = cKDTree(pts).query_pairs(np.sqrt(6 / (np.pi * m)),
                                     output_type='ndarray')
    g = rng.uniform(5, 12, len(pairs))
    conns = [(int(i) + k * m, int(j) + k * m, float(w))
             for k in range(K) for (i, j), w in zip(pairs, g)]
    left = [int(c) + k * m for k in range(K)
            for c in np.nonzero(pts[:, 0] < 0.1)[0]]
    gna = 16 * np.exp(rng.uniform(np.log(.5), np.log(2), m * K))
    probes = [int(np.argmin(np.abs(pts[:, 0] - 0.4))) + k * m
              for k in range(K)]

    def run():
        s = myokit.SimulationOpenCL(
            myokit.load_model('example'), myokit.pacing.blocktrain(1000, 2),
            ncells=m * K, precision=64, rl=True)
        s.set_step_size(0.01)
        s.set_connections(conns)
        s.set_paced_cell_list(left)
        s.set_field('ina.gNa', gna)
        adj = TissueAdjoint(s, probes, checkpoint_every=100)
        target = np.full((adj.n_steps(4.0), K), -60.0)
        loss, grads, _, _ = adj.loss_and_gradient(4.0, target)
        return s, loss, np.array(grads['ina.gNa'])

    s1, l1, g1 = run()
    assert s1._copies == K
    orig = simulation._graph_copies
    simulation._graph_copies = lambda *a, **k: 1
    try:
        s2, l2, g2 = run()
    finally:
        simulation._graph_copies = orig
    assert s2._copies == 1
    print('graph copies adjoint: loss %.17g vs %.17g, max|dg| %.3g'
          % (l1, l2, float(np.max(np.abs(g1 - g2)))), flush=True)
    assert l1 == l2
    np.testing.assert_array_equal(g1, g2)


@pytest.mark.parametrize('K', [23, 599])
def test_adjoint_overlapped_recompute_bit_identical_fp64(K):
    # On a small domain the grid path recomputes each earlier segment on a
    # second stream while the later segment's reverse steps run, the two
    # segments' states in two trajectory buffers. The kernels and their
    # inputs are those of the in-line order, so loss and gradients must be
    # identical bit for bit. 600 steps: K = 23 gives 27 segments (the last
    # one shorter), K = 599 two (the last of one step).
    res = []
    for overlap in (1, 0):
        s = cable(n=200)
        s.set_engine_option('driver', 2)
        s.set_engine_option('adj_overlap', overlap)
        adj = TissueAdjoint(s, [3, 77, 150, 199], checkpoint_every=K)
        assert adj.n_steps(3.0) == 600
        target = np.full((600, 4), -60.0)
        loss, g, dgx, _ = adj.loss_and_gradient(3.0, target)
        kind, launches = adj.reverse_kernel()
        print('K', K, 'overlap', overlap, 'reverse', kind, launches, 'loss',
              loss, 'dgx', dgx, 'phases', adj.timing(), flush=True)
        assert s.engine_driver() == 'grid'
        assert kind == ('step-overlap' if overlap else 'step')
        assert launches == 601
        res.append((loss, {k: np.array(v) for k, v in g.items()}, dgx))
    (l1, g1, x1), (l2, g2, x2) = res
    assert l1 == l2 and x1 == x2
    for name in ('ina.gNa', 'ica.gCa'):
        assert np.array_equal(g1[name], g2[name]), name
