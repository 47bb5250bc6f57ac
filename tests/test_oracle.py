"""
CPU tests that pin the oracle itself: against the reference's Simulation1d
golden runs, against invariants of the coupling (uniform tissue, symmetry,
zero net diffusion current, connections equal to the stencil), and the
portable math header against the C library.
"""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

import myokit_amd as myokit
from oracle.cgen import OracleSimulation

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, 'golden')
ROOT = os.path.dirname(HERE)
MODELS = {'lr1991': 'example',
          'br1977': os.path.join(HERE, 'data', 'beeler-1977-model.mmt')}


SIM1D = [('sim1d_lr1991_fe', 'lr1991'), ('sim1d_lr1991_rl', 'lr1991'),
         ('sim1d_br1977_fe', 'br1977'), ('sim1d_lr1991_cfg1', 'lr1991'),
         ('sim1d_br1977_reftest', 'br1977')]


def sim1d_oracle(name, model):
    """The oracle configured as the Simulation1d golden run ``name``."""
    g = np.load(os.path.join(GOLD, name + '.npz'))
    n, npaced, gx, rl, dt, dur, li = g['config']
    m = myokit.load_model(MODELS[model])
    if np.isfinite(g['amplitude']):
        m.get('membrane.i_stim.amplitude').set_rhs(float(g['amplitude']))
    p = myokit.Protocol()
    p.schedule(*g['event'])
    o = OracleSimulation(m, p, ncells=int(n), rl=bool(rl))
    o.gx = gx
    o.set_paced_cells(int(npaced))
    o.dt = dt
    return o, g


def sim1d_log_vars(g):
    return ['membrane.V', 'membrane.i_diff', str(g['inter_name'])]


@pytest.mark.parametrize('name,model', SIM1D)
def test_oracle_matches_reference_simulation1d(name, model):
    # Every logged quantity of a row: V, the diffusion current, an
    # intermediary, the pacing level and the time (the reference's own
    # comparison, test_simulation_opencl_vs_sim1d.py:28-136)
    o, g = sim1d_oracle(name, model)
    dur, li = g['config'][5:7]
    lv = sim1d_log_vars(g)
    r = o.run(dur, log_vars=lv, log_interval=li)
    np.testing.assert_array_equal(r['time'], g['time'])
    np.testing.assert_array_equal(r['pace'], g['pace'])
    assert np.max(g['pace']) == 1 and np.min(g['pace']) == 0
    # The reference pins SimulationOpenCL to Simulation1d within 1e-13
    # (test_simulation_opencl_vs_sim1d.py:118-136); the remaining
    # difference is the last bits of exp/log of the two math libraries
    assert np.max(np.abs(r['membrane.V'] - g['V'])) < 1e-11
    assert np.max(np.abs(r['membrane.i_diff'] - g['i_diff'])) < 1e-11
    assert np.max(np.abs(r[lv[2]] - g['inter'])) < 1e-11
    assert np.max(np.abs(g['i_diff'])) > 1 and np.max(np.abs(g['inter'])) > 0
    assert np.max(np.abs(o.state - g['final_state'])) < 1e-11


@pytest.mark.parametrize('name,model', SIM1D[:1] + SIM1D[2:3])
def test_golden_logs_catch_values_logged_after_the_update(name, model):
    # A planted mistake in the oracle (each log row evaluated from the
    # state after the row's step instead of before it) must fail the
    # golden comparison of V and of the intermediary
    o, g = sim1d_oracle(name, model)
    o.log_after_update = True
    dur, li = g['config'][5:7]
    lv = sim1d_log_vars(g)
    r = o.run(dur, log_vars=lv, log_interval=li)
    assert np.max(np.abs(r['membrane.V'] - g['V'])) > 1e-6
    assert np.max(np.abs(r[lv[2]] - g['inter'])) > 1e-8


def test_uniform_unpaced_sheet_equals_single_cell():
    # With identical cells and no stimulus, every diffusion current is
    # exactly zero: each cell follows the single-cell trajectory
    m = myokit.load_model('example')
    p = myokit.pacing.blocktrain(1000, 2)
    o = OracleSimulation(m, p, ncells=(6, 5), rl=True)
    o.paced = []
    o.dt = 0.01
    r = o.run(5, log_vars=['membrane.V'], log_interval=0.5)
    V = r['membrane.V']
    assert np.all(V == V[:, :1])
    u = OracleSimulation(m, p, ncells=1, diffusion=False, rl=True)
    u.dt = 0.01
    u.paced = None
    # uncoupled, but pace off: use a protocol-free run
    u.protocol = None
    ru = u.run(5, log_vars=['membrane.V'], log_interval=0.5)
    np.testing.assert_array_equal(V[:, 0], ru['membrane.V'][:, 0])


def test_column_paced_sheet_rows_identical():
    # Pacing a whole column: the y-coupling terms vanish, every row is the
    # same 1D cable
    m = myokit.load_model('example')
    p = myokit.pacing.blocktrain(1000, 2)
    o = OracleSimulation(m, p, ncells=(12, 4), rl=True)
    o.set_paced_cells(2, 4, 0, 0)
    o.gx, o.gy = 9.5, 7.0
    o.dt = 0.01
    r = o.run(10, log_vars=['membrane.V'], log_interval=0.5)
    V = r['membrane.V'].reshape(-1, 4, 12)
    for row in range(1, 4):
        np.testing.assert_array_equal(V[:, row], V[:, 0])
    c = OracleSimulation(m, p, ncells=12, rl=True)
    c.set_paced_cells(2)
    c.gx = 9.5
    c.dt = 0.01
    rc = c.run(10, log_vars=['membrane.V'], log_interval=0.5)
    np.testing.assert_array_equal(V[:, 0], rc['membrane.V'])


@pytest.mark.parametrize('shape', [20, (7, 5)])
def test_net_diffusion_current_is_zero(shape):
    m = myokit.load_model('example')
    p = myokit.pacing.blocktrain(1000, 2)
    o = OracleSimulation(m, p, ncells=shape, rl=True)
    o.dt = 0.01
    r = o.run(4, log_vars=['membrane.i_diff'], log_interval=0.25)
    i = r['membrane.i_diff']
    scale = np.max(np.abs(i), axis=1) + 1e-300
    assert np.max(np.abs(i.sum(axis=1)) / scale) < 1e-12
    assert np.max(scale) > 1


def test_connections_chain_equals_stencil():
    m = myokit.load_model('example')
    p = myokit.pacing.blocktrain(1000, 2)
    a = OracleSimulation(m, p, ncells=30)
    a.gx = 9.5
    a.set_paced_cells(3)
    b = OracleSimulation(m, p, ncells=30)
    b.connections = [(i, i + 1, 9.5) for i in range(29)]
    b.paced = [0, 1, 2]
    ra = a.run(10, log_vars=['membrane.V'], log_interval=0.5)
    rb = b.run(10, log_vars=['membrane.V'], log_interval=0.5)
    assert np.max(np.abs(ra['membrane.V'] - rb['membrane.V'])) <= 1e-9


def test_portable_math_against_libm():
    src = os.path.join(HERE, 'mk_math_check.c')
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, 'check')
        subprocess.run(
            ['gcc', '-O2', '-ffp-contract=off', '-I',
             os.path.join(ROOT, 'myokit_amd', 'csrc'), '-o', exe, src,
             '-lm'], check=True)
        out = subprocess.run([exe], capture_output=True, text=True,
                             check=True).stdout
    res = dict(line.split() for line in out.strip().splitlines())
    ulps = {k: float(v) for k, v in res.items()}
    assert ulps['exp'] <= 1.0
    assert ulps['log'] <= 1.0
    assert ulps['pow_int'] <= 2.0
    assert ulps['pow'] <= 128.0
    assert ulps['sin'] <= 2.0 and ulps['cos'] <= 2.0
    assert ulps['atan'] <= 8.0
    assert ulps['special'] == 0.0


def _sheet(shape, gx, gy, paced, duration=12.0, log_vars=('membrane.V',),
           field=None, fields=None):
    m = myokit.load_model('example')
    p = myokit.pacing.blocktrain(1000, 2)
    o = OracleSimulation(m, p, ncells=shape, rl=True)
    o.set_paced_cells(*paced)
    o.gx, o.gy = gx, gy
    o.dt = 0.01
    if field is not None:
        o.gx_field, o.gy_field = field
    for k, v in (fields or {}).items():
        o.fields[k] = v
    return o.run(duration, log_vars=list(log_vars), log_interval=0.5)


def test_transposed_sheet_with_swapped_conductances():
    # A (nx, ny) sheet with (gx, gy) and a corner stimulus equals the
    # transposed (ny, nx) sheet with (gy, gx) and the transposed stimulus,
    # bit for bit (each stencil term is the same product, and the two
    # directions' terms are added in swapped order: a + b = b + a). A y
    # term that used gx, a wrong row stride or a transposed neighbour
    # index fails this.
    nx, ny = 13, 9
    a = _sheet((nx, ny), 9.5, 4.0, (4, 3, 0, 0),
               log_vars=('membrane.V', 'membrane.i_diff'))
    b = _sheet((ny, nx), 4.0, 9.5, (3, 4, 0, 0),
               log_vars=('membrane.V', 'membrane.i_diff'))
    for v in ('membrane.V', 'membrane.i_diff'):
        A = a[v].reshape(-1, ny, nx)
        B = b[v].reshape(-1, nx, ny).transpose(0, 2, 1)
        np.testing.assert_array_equal(A, B)
    # the wave is two-dimensional: both directions carry current
    A = a['membrane.V'].reshape(-1, ny, nx)
    assert np.max(np.abs(A[:, 0, :] - A[:, -1, :])) > 10
    assert np.max(np.abs(A[:, :, 0] - A[:, :, -1])) > 10


def test_uniform_conductance_field_equals_stencil():
    # A conductance field holding gx on every x connection and gy on every
    # y connection is the homogeneous stencil (diff_hetero and diff_step of
    # openclsim.cl:384-487 describe the same Laplacian), under a corner
    # stimulus that makes every term nonzero. The two sum in a different
    # order, so they agree to rounding, not bit for bit.
    nx, ny = 12, 10
    gx, gy = 9.5, 4.0
    field = (np.full(ny * (nx - 1), gx), np.full((ny - 1) * nx, gy))
    a = _sheet((nx, ny), gx, gy, (3, 4, 0, 0))
    b = _sheet((nx, ny), 0.0, 0.0, (3, 4, 0, 0), field=field)
    d = np.max(np.abs(a['membrane.V'] - b['membrane.V']))
    assert d < 1e-8
    # and a field with the two directions swapped is not (gx != gy)
    c = _sheet((nx, ny), 0.0, 0.0, (3, 4, 0, 0),
               field=(np.full(ny * (nx - 1), gy), np.full((ny - 1) * nx, gx)))
    assert np.max(np.abs(a['membrane.V'] - c['membrane.V'])) > 1


def test_field_at_the_constants_value_equals_no_field():
    # set_field with every cell at the constant's own value changes nothing
    # (openclsim.py:1082-1089 substitutes the field for the constant): here
    # for a constant used directly (ina.gNa = 16) and one that other constants
    # derive from (cell.K_o: ik.gK, ik.E, ik1.gK1, ik1.E, ...), whose
    # dependants the oracle then computes per cell in C instead of folding
    # them in Python
    shape = (8, 6)
    n = 48
    base = _sheet(shape, 9.5, 4.0, (3, 2, 0, 0),
                  log_vars=('membrane.V', 'ik.IK'))
    f = _sheet(shape, 9.5, 4.0, (3, 2, 0, 0), log_vars=('membrane.V', 'ik.IK'),
               fields={'ina.gNa': np.full(n, 16.0), 'cell.K_o': np.full(n, 5.4)})
    for v in ('membrane.V', 'ik.IK'):
        assert np.max(np.abs(base[v] - f[v])) <= 1e-12 * np.max(np.abs(base[v]))
    # a field 1% off the constant is visible
    g = _sheet(shape, 9.5, 4.0, (3, 2, 0, 0), log_vars=('membrane.V',),
               fields={'cell.K_o': np.full(n, 5.4 * 1.01)})
    assert np.max(np.abs(base['membrane.V'] - g['membrane.V'])) > 1e-3
