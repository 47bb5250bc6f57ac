"""
CPU tests of the SimulationOpenCL front end: argument checks and messages of
the reference (myokit/tests/test_simulation_opencl.py), state indexing,
paced-cell geometry, neighbours, log preparation and keys. None of these
touch the device.
"""
import numpy as np
import pytest

import myokit_amd as myokit


@pytest.fixture(scope='module')
def model():
    return myokit.load_model('example')


def test_creation_errors(model):
    # test_simulation_opencl.py:530-600
    with pytest.raises(ValueError, match='scalar or a tuple'):
        myokit.SimulationOpenCL(model, ncells=(1, 2, 3))
    with pytest.raises(ValueError, match='at least 1'):
        myokit.SimulationOpenCL(model, ncells=0)
    with pytest.raises(ValueError, match='single and double'):
        myokit.SimulationOpenCL(model, precision=16)
    m = model.clone()
    m.get('membrane.V').set_label(None)
    with pytest.raises(ValueError, match='membrane_potential'):
        myokit.SimulationOpenCL(m)
    myokit.SimulationOpenCL(m, diffusion=False)


def test_conductance(model):
    s = myokit.SimulationOpenCL(model, ncells=5)
    assert s.conductance() == 10
    s.set_conductance(3, 4)
    assert s.conductance() == 3
    s = myokit.SimulationOpenCL(model, ncells=(5, 4))
    assert s.conductance() == (10, 5)
    with pytest.raises(ValueError, match='Invalid conductance gx'):
        s.set_conductance(-1, 1)
    with pytest.raises(ValueError, match='Invalid conductance gy'):
        s.set_conductance(1, -1)
    s = myokit.SimulationOpenCL(model, ncells=5, diffusion=False)
    with pytest.raises(RuntimeError, match='diffusion is disabled'):
        s.conductance()
    with pytest.raises(RuntimeError, match='diffusion is disabled'):
        s.set_paced_cells()


def test_conductance_field_checks(model):
    s = myokit.SimulationOpenCL(model, ncells=(4, 3))
    with pytest.raises(ValueError, match='dimensions'):
        s.set_conductance_field(np.ones((3, 4)), np.ones((2, 4)))
    with pytest.raises(ValueError, match='must be set'):
        s.set_conductance_field(np.ones((3, 3)))
    with pytest.raises(ValueError, match='negative'):
        s.set_conductance_field(-np.ones((3, 3)), np.ones((2, 4)))
    s.set_conductance_field(np.ones((3, 3)), np.ones((2, 4)))
    assert s.conductance() is None
    s = myokit.SimulationOpenCL(model, ncells=4)
    with pytest.raises(ValueError, match='length 3'):
        s.set_conductance_field([1, 2])
    with pytest.raises(ValueError, match='must be None'):
        s.set_conductance_field([1, 2, 3], [1])


def test_connection_checks(model):
    # test_simulation_opencl.py:324-392
    s = myokit.SimulationOpenCL(model, ncells=4)
    with pytest.raises(ValueError, match='cannot be None'):
        s.set_connections(None)
    with pytest.raises(ValueError, match='3-tuples'):
        s.set_connections([(0, 1)])
    with pytest.raises(ValueError, match='Invalid connection'):
        s.set_connections([(0, 4, 1)])
    with pytest.raises(ValueError, match='Invalid connection'):
        s.set_connections([(1, 1, 1)])
    with pytest.raises(ValueError, match='Duplicate connection'):
        s.set_connections([(0, 1, 1), (1, 0, 2)])
    with pytest.raises(ValueError, match='Invalid conductance'):
        s.set_connections([(0, 1, -1)])
    s.set_connections([(2, 1, 1.5), (0, 1, 2)])
    assert s._connections == [(1, 2, 1.5), (0, 1, 2.0)]
    assert sorted(s.neighbors(1)) == [0, 2]
    s2 = myokit.SimulationOpenCL(model, ncells=(2, 2))
    with pytest.raises(RuntimeError, match='1d mode'):
        s2.set_connections([(0, 1, 1)])


def test_fields(model):
    s = myokit.SimulationOpenCL(model, ncells=(3, 2))
    with pytest.raises(ValueError, match='Bound values'):
        s.set_field('engine.pace', np.zeros((2, 3)))
    with pytest.raises(ValueError, match='Only constants'):
        s.set_field('ina.INa', np.zeros((2, 3)))
    with pytest.raises(ValueError, match='dimensions'):
        s.set_field('ina.gNa', np.zeros((3, 2)))
    s.set_field('ina.gNa', np.ones((2, 3)))
    s.set_field('ik.IK.gK', np.ones((2, 3)))
    s.remove_field('ina.gNa')
    assert list(f.qname() for f in s._fields) == ['ik.IK.gK']


def test_paced_cells_geometry(model):
    # test_simulation_opencl.py:1002-1219
    s = myokit.SimulationOpenCL(model, ncells=(10, 8))
    s.set_paced_cells(3, 2, 1, 1)
    assert s._paced_cells == (3, 2, 1, 1)
    assert s.is_paced(1, 1) and s.is_paced(3, 2) and not s.is_paced(4, 2)
    s.set_paced_cells(-3, -2, 10, 8)
    assert s._paced_cells == (3, 2, 7, 6)
    s.set_paced_cells(2, 2, -2, -2)
    assert s._paced_cells == (2, 2, 8, 6)
    s.set_paced_cell_list([(0, 0), (9, 7)])
    assert s._paced_cells == [0, 79]
    assert s.is_paced(9, 7) and not s.is_paced(8, 7)
    with pytest.raises(IndexError):
        s.set_paced_cell_list([(10, 0)])
    with pytest.raises(ValueError, match='No y-coordinate'):
        s.is_paced(1)
    s = myokit.SimulationOpenCL(model, ncells=10)
    s.set_paced_cells(3, 4, 5, 6)
    assert s._paced_cells == (3, 1, 5, 0)
    with pytest.raises(ValueError, match='Y-coordinate'):
        s.is_paced(1, 1)


def test_neighbors(model):
    s = myokit.SimulationOpenCL(model, ncells=(3, 3))
    assert s.neighbors(0, 0) == [(1, 0), (0, 1)]
    assert s.neighbors(1, 1) == [(0, 1), (2, 1), (1, 0), (1, 2)]
    s = myokit.SimulationOpenCL(model, ncells=3)
    assert s.neighbors(0) == [1] and s.neighbors(1) == [0, 2]
    with pytest.raises(IndexError):
        s.neighbors(3)


def test_state_indexing_x_first(model):
    # test_simulation_opencl.py:1221-1410
    s = myokit.SimulationOpenCL(model, ncells=(3, 2))
    n = model.count_states()
    full = np.arange(6 * n, dtype=float)
    s.set_state(full)
    assert s.state(1, 1) == list(full[4 * n:5 * n])
    s.set_state(list(range(n)), 2, 0)
    assert s.state(2, 0) == list(range(n))
    with pytest.raises(ValueError, match='argument x'):
        s.set_state(full, 1)
    with pytest.raises(ValueError, match='same size'):
        s.set_state([1, 2])
    with pytest.raises(IndexError):
        s.state(3, 0)
    s.set_default_state(model.state())
    assert s.default_state(0, 1) == model.state()
    assert s.shape() == (2, 3)
    with pytest.raises(ValueError, match='negative'):
        s.run(-1)


def test_prepare_log_keys(model):
    m = model
    d = myokit.prepare_log(['engine.time', 'membrane.V'], m, dims=(3, 2),
                           global_vars=['engine.time', 'engine.pace'])
    assert list(d.keys()) == ['engine.time', '0.0.membrane.V',
                              '1.0.membrane.V', '2.0.membrane.V',
                              '0.1.membrane.V', '1.1.membrane.V',
                              '2.1.membrane.V']
    assert d.time_key() == 'engine.time'
    d = myokit.prepare_log(myokit.LOG_STATE + myokit.LOG_BOUND, m, dims=(2,),
                           global_vars=['engine.time', 'engine.pace'])
    assert '1.ica.Ca_i' in d and 'engine.pace' in d and '0.membrane.i_diff' in d
    with pytest.raises(ValueError, match='Unknown variable'):
        myokit.prepare_log(['x.y'], m, dims=(2,))
    with pytest.raises(ValueError, match='constants'):
        myokit.prepare_log(['ina.gNa'], m, dims=(2,))
    with pytest.raises(ValueError, match='Invalid index'):
        myokit.prepare_log(['5.membrane.V'], m, dims=(2,))
    with pytest.raises(ValueError, match='global'):
        myokit.prepare_log(['0.engine.time'], m, dims=(2,),
                           global_vars=['engine.time'])
    assert myokit.split_key('12.3.membrane.V') == ('12.3.', 'membrane.V')
    assert myokit.split_key('membrane.V') == ('', 'membrane.V')


def test_monodomain_conductance(model):
    s = myokit.SimulationOpenCL(model, ncells=3)
    # g = D * A * k / ((k + 1) * chi * dx^2)
    assert s.monodomain_conductance(2, 3, 4, 0.5, 1) == pytest.approx(
        4 * 3 / (4 * 2 * 0.25))


def test_connections_as_array(model):
    # An (E, 3) array takes the vectorised path of set_connections (graphs
    # of millions of connections); it stores what the list form stores and
    # raises the same errors with the same messages.
    s = myokit.SimulationOpenCL(model, ncells=10)
    conns = [(0, 1, 9.5), (2, 1, 3.0), (5, 6, 1.0)]
    s.set_connections(conns)
    ref = np.array(s._connections, dtype=float)
    s.set_connections(np.array(conns))
    np.testing.assert_array_equal(np.asarray(s._connections), ref)
    assert s.neighbors(1) == [0, 2]
    for bad, msg in (([(0, 0, 1.0)], 'Invalid connection: \\(0, 0'),
                     ([(0, 10, 1.0)], 'Invalid connection'),
                     ([(0, 1, 1.0), (1, 0, 2.0)], 'Duplicate connection'),
                     ([(0, 1, -1.0)], 'Invalid conductance')):
        for form in (bad, np.array(bad)):
            with pytest.raises(ValueError, match=msg):
                s.set_connections(form)
    with pytest.raises(ValueError, match='3-tuples'):
        s.set_connections(np.zeros((4, 2)))


def test_kernel_source_cache(model):
    # A repeated run reuses its generated kernel source; set_constant (the
    # generator inlines constants) and a change of fields or tuning make a
    # new one.
    s = myokit.SimulationOpenCL(model, ncells=8, rl=True)
    q = [s._model.get('membrane.V')]
    a, _ = s._kernel_source(q)
    b, _ = s._kernel_source(q)
    assert a is b
    s.set_constant('ina.gNa', 17.0)
    c, _ = s._kernel_source(q)
    assert c is not a and c != a
    assert s._kernel_source(q)[0] is c
    s.set_field('ina.gNa', np.ones(8))
    assert s._kernel_source(q)[0] != c
    s.set_tuning(sheet_cpt=2)
    d, _ = s._kernel_source(q)
    assert s._kernel_source(q, adjoint=True)[0] != d


def test_threads_per_workgroup_follow_the_layout(model):
    # Unless set with set_tuning, the block layout gets one thread per cell
    # of its largest domain (multiple of 64, at most 1024); other layouts
    # 256.
    s = myokit.SimulationOpenCL(model, ncells=200, precision=64)
    assert s.tuning()['block_threads'] == 256          # one 200-cell cable
    s = myokit.SimulationOpenCL(model, ncells=384)
    assert s.tuning()['block_threads'] == 384
    s = myokit.SimulationOpenCL(model, ncells=1000)
    assert s.tuning()['block_threads'] == 256          # grid layout
    s = myokit.SimulationOpenCL(model, ncells=(24, 16))
    assert s.tuning()['block_threads'] == 384
    s = myokit.SimulationOpenCL(model, ncells=(1024, 1024))
    assert s.tuning()['block_threads'] == 256
    s = myokit.SimulationOpenCL(model, ncells=4096, diffusion=False)
    assert s.tuning()['block_threads'] == 256
    # cables of 50 .. 2000 cells as one graph: largest component 2000
    lengths, rows, off = [50, 2000, 700], [], 0
    for L in lengths:
        i = np.arange(off, off + L - 1)
        rows.append(np.stack([i, i + 1, np.full(L - 1, 9.5)], axis=1))
        off += L
    s = myokit.SimulationOpenCL(model, ncells=off)
    s.set_connections(np.concatenate(rows))
    assert s._graph_domains() == (2000, 3)
    assert s.tuning()['block_threads'] == 1024
    s64 = myokit.SimulationOpenCL(model, ncells=off, precision=64)
    s64.set_connections(np.concatenate(rows))
    assert s64.tuning()['block_threads'] == 1024      # 2000 <= 2048
    s.set_connections([(0, 1, 1.0)])                  # components 2, 1, ..
    assert s._graph_domains() == (2, off - 1)
    assert s.tuning()['block_threads'] == 64
    s.set_tuning(block_threads=512)
    assert s.tuning()['block_threads'] == 512


def test_graph_cell_order(model):
    # set_connections graphs live on the device in reverse Cuthill-McKee
    # order: a permutation and its inverse, recomputed for a new graph and
    # dropped with the graph.
    rng = np.random.default_rng(4)
    n = 300
    s = myokit.SimulationOpenCL(model, ncells=n)
    assert s._cell_order() == (None, None)
    pts = rng.uniform(0, 1, (n, 2))
    from scipy.spatial import cKDTree
    pairs = cKDTree(pts).query_pairs(0.12, output_type='ndarray')
    rows = np.concatenate([pairs, np.full((len(pairs), 1), 5.0)], axis=1)
    s.set_connections(rows)
    perm, iperm = s._cell_order()
    assert sorted(perm) == list(range(n))
    np.testing.assert_array_equal(iperm[perm], np.arange(n))
    i, j = pairs[:, 0], pairs[:, 1]
    assert np.mean(np.abs(iperm[i] - iperm[j])) < 0.5 * np.mean(np.abs(i - j))
    assert s._cell_order()[0] is perm                 # cached
    s.set_connections([(0, 1, 1.0), (1, 2, 1.0)])
    p2, _ = s._cell_order()
    assert p2 is None or p2 is not perm
    s.set_conductance(9.5)
    assert s._cell_order() == (None, None)


def test_graph_copies_interleaved(model):
    # A batch of identical graph copies (cells k m .. (k + 1) m - 1 for
    # copy k) is held with the copies interleaved: device cell p K + k is
    # copy k's cell at position p of the first copy's reverse Cuthill-McKee
    # order. Copies of at most a workgroup's worth of cells, sets of chains
    # and blocks that differ keep the plain order.
    from myokit_amd.simulation import _graph_copies
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(8)
    m, K = 5000, 6
    pts = rng.uniform(0, 1, (m, 2))
    pairs = cKDTree(pts).query_pairs(np.sqrt(6 / (np.pi * m)),
                                     output_type='ndarray')
    rows = np.concatenate([
        np.concatenate([pairs + k * m,
                        rng.uniform(5, 12, (len(pairs), 1))], axis=1)
        for k in range(K)])
    i, j = rows[:, 0].astype(np.int64), rows[:, 1].astype(np.int64)
    assert _graph_copies(i, j, m * K, 4096) == K
    # the copies' connections in another order still count
    sh = rng.permutation(len(i))
    assert _graph_copies(i[sh], j[sh], m * K, 4096) == K
    assert _graph_copies(i, j, m * K, 8192) == 1        # one workgroup each
    assert _graph_copies(i[:-1], j[:-1], m * K, 4096) == 1  # a copy differs
    chain = np.arange(m * K - 1)
    keep = (chain + 1) % m != 0
    assert _graph_copies(chain[keep], chain[keep] + 1, m * K, 4096) == 1
    s = myokit.SimulationOpenCL(model, ncells=m * K)
    s.set_connections(rows)
    perm, iperm = s._cell_order()
    assert sorted(perm) == list(range(m * K))
    np.testing.assert_array_equal(iperm[perm], np.arange(m * K))
    # cell p of every copy side by side, the same position in each copy
    P = perm.reshape(-1, K)
    assert np.all(P - P[:, :1] == np.arange(K)[None, :] * m)
    # and neighbours close together within a copy
    d = np.abs(iperm[i] - iperm[j]) // K
    assert np.mean(d) < 0.5 * np.mean(np.abs(i - j))


def test_adjoint_probe_checks(model):
    # The loss keeps one probe per cell: a cell listed twice is an error
    # rather than a silently dropped loss term (the library rejects it too)
    from myokit_amd.adjoint import TissueAdjoint
    s = myokit.SimulationOpenCL(model, ncells=10, precision=64)
    TissueAdjoint(s, [0, 3, 9])
    with pytest.raises(ValueError, match='distinct'):
        TissueAdjoint(s, [0, 3, 3])
    with pytest.raises(ValueError, match='out of range'):
        TissueAdjoint(s, [0, 10])
    s = myokit.SimulationOpenCL(model, ncells=(4, 3), precision=64)
    with pytest.raises(ValueError, match='distinct'):
        TissueAdjoint(s, [(1, 2), (1, 2)])


def test_library_releases_handles_at_exit():
    # The atexit hook frees every engine, program and pooled page-locked
    # block before the C runtime's exit handlers run (and before a profiler
    # attached to the HIP runtime shuts down); later finalizers make no
    # library calls.
    import subprocess
    import sys
    code = (
        'import atexit, myokit_amd._native as n\n'
        'n.lib()\n'
        'class E: pass\n'
        'e = E(); e.h = 12345\n'
        'calls = []\n'
        'n.track(e, calls.append)\n'
        'n._shutdown()\n'
        'assert calls == [12345] and e.h is None and n.closed\n'
        'try:\n'
        '    n.lib()\n'
        'except n.DeviceError:\n'
        '    print("closed")\n')
    r = subprocess.run([sys.executable, '-c', code], capture_output=True,
                       text=True, cwd=__import__('os').path.dirname(
                           __import__('os').path.dirname(__file__)))
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == 'closed'
