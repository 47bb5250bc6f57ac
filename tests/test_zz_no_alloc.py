"""
GPU test of mkh_adj_trace_static, the variant of mkh_adj_trace that
allocates nothing (its buffers live in a static device block of
MKH_STATIC_SCRATCH_BYTES): called through the C ABI, its trace must equal
mkh_adj_trace's bit for bit, on two calls in a row, the second with a
smaller input (a smaller simulation, fewer probes, a shorter run), which
also finds the block holding the first call's contents. A run whose
buffers exceed the block is refused before anything is launched.
"""
import ctypes

import numpy as np
import pytest

import myokit_amd as myokit
from myokit_amd import _native
from myokit_amd.adjoint import TissueAdjoint

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 32 << 20     # MKH_STATIC_SCRATCH_BYTES


def trace_static(adj, duration):
    """adj.trace(duration) through mkh_adj_trace_static."""
    sim = adj.sim
    eng = adj._engine()
    ev, nev = adj._events()
    n = adj.n_steps(duration)
    out = np.full((n, len(adj.probes)), np.nan)
    probes = adj._device_probes()
    code = _native.lib().mkh_adj_trace_static(
        eng.h, sim._time, sim._time + duration, sim._step_size,
        _native.ptr(ev, ctypes.c_double), nev,
        _native.ptr(probes, ctypes.c_longlong), len(probes),
        _native.ptr(out, ctypes.c_double), n)
    return code, out


def _sheet(shape, precision, fields=False):
    s = myokit.SimulationOpenCL(
        myokit.load_model('example'), myokit.pacing.blocktrain(1000, 2),
        ncells=shape, rl=True, precision=precision)
    s.set_step_size(0.01)
    s.set_conductance(9.5, 6.0)
    s.set_paced_cells(4, 3, 0, 0)
    if fields:
        rng = np.random.default_rng(2)
        s.set_field('ina.gNa', 16 * rng.uniform(0.7, 1.3, size=shape[::-1]))
    return s


def test_static_trace_equals_trace_twice():
    # First call: a 100 x 70 single-precision sheet (persistent sheet
    # driver) with a field, 30 probes, 600 steps
    s = _sheet((100, 70), myokit.SINGLE_PRECISION, fields=True)
    rng = np.random.default_rng(5)
    probes = [(int(x), int(y)) for x, y in zip(rng.integers(0, 100, 30),
                                               rng.integers(0, 70, 30))]
    probes = list(dict.fromkeys(probes))
    adj = TissueAdjoint(s, probes)
    ref = adj.trace(6.0)
    code, got = trace_static(adj, 6.0)
    print('static trace 1:', got.shape, 'driver', s.engine_driver(),
          flush=True)
    assert code == 0, _native.last_error()
    np.testing.assert_array_equal(got, ref)
    assert np.max(ref) > 0 > np.min(ref)      # a wave crossed some probes
    # Second call, smaller: a 40-cell double-precision cable (grid driver
    # forced) with 3 probes and 150 steps; the block still holds the first
    # call's contents
    c = myokit.SimulationOpenCL(
        myokit.load_model('example'), myokit.pacing.blocktrain(1000, 2),
        ncells=40, precision=myokit.DOUBLE_PRECISION)
    c.set_conductance(9.5)
    c.set_paced_cells(3)
    c.set_step_size(0.01)
    c.set_tuning(driver='grid')
    adj2 = TissueAdjoint(c, [0, 17, 39])
    ref2 = adj2.trace(1.5)
    code, got2 = trace_static(adj2, 1.5)
    print('static trace 2:', got2.shape, 'driver', c.engine_driver(),
          flush=True)
    assert code == 0, _native.last_error()
    np.testing.assert_array_equal(got2, ref2)
    # and the first simulation again, now after a smaller call
    code, again = trace_static(adj, 6.0)
    assert code == 0, _native.last_error()
    np.testing.assert_array_equal(again, ref)
    # the state and time are left unchanged, as by mkh_adj_trace
    assert s.time() == 0 and c.time() == 0


def test_static_trace_refuses_runs_beyond_the_block():
    # 700 x 700 cells in double precision: two state copies alone are
    # 2 * 8 * 490000 * 8 bytes, beyond the 32 MiB block
    s = _sheet((700, 700), myokit.DOUBLE_PRECISION)
    adj = TissueAdjoint(s, [(1, 1)])
    code, got = trace_static(adj, 0.05)
    assert code == -1
    msg = _native.last_error()
    print('refused:', msg, flush=True)
    assert 'static scratch' in msg and str(SCRATCH_BYTES) in msg
    assert np.all(np.isnan(got))
