"""
The split cell step (``mk_cell_a`` then ``mk_cell_b``, used by the
persistent sheet and chain drivers to overlap the neighbour exchange with
arithmetic) against the whole step ``mk_cell`` that every other driver
runs: compiled for the host with g++ (host stand-ins for the two device
builtins, the same in both), the states and logged quantities must be
equal bit for bit, in double and single precision, for LR1991 (forward
Euler and Rush-Larsen, with a field and logged intermediaries, the
diffusion current among them) and for a model in which the diffusion
current also feeds an intermediary and a second state.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import myokit_amd as myokit
from myokit_amd import _kernelgen
from myokit_amd.simulation import _prepare_model

CSRC = os.path.join(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__))), 'myokit_amd', 'csrc')

PRELUDE = '''
#include <math.h>
#include "mk_math.h"
#define __device__ static
#define __forceinline__ inline
#define __builtin_amdgcn_exp2f(x) exp2f(x)
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))
typedef %s Real;
'''

WRAPPER = '''
extern "C" void whole(Real* y, Real t, Real dt, Real pace, Real idiff,
                      const Real* fv, Real* q)
{
    mk_cell<true>(y, t, dt, pace, idiff, fv, q);
}
extern "C" void split(Real* y, Real t, Real dt, Real pace, Real idiff,
                      const Real* fv, Real* q)
{
    Real carry[MK_NCARRY];
    mk_cell_a<true>(y, t, dt, pace, fv, q, carry);
    mk_cell_b<true>(y, t, dt, pace, idiff, fv, q, carry);
}
extern "C" int qpart(int k) { return mk_qpart[k]; }
'''

SIDE_MODEL = '''
[[model]]
c.v = -80
c.w = 0.1
c.u = 0.5

[engine]
time = 0 bind time
pace = 0 bind pace

[c]
dot(v) = -(ion + stim + s)
    label membrane_potential
ion = 0.1 * (v + 80) * w
stim = engine.pace * -10
idiff = 0 bind diffusion_current
s = idiff * 1.5
dot(w) = (exp(-v / 20) - w) / 5
dot(u) = s * 0.01 - u
'''


def _build(tmp_path, cm, prec, tag):
    ct = 'double' if prec == 64 else 'float'
    src = PRELUDE % ct
    if prec == 32:
        src += _kernelgen._FLOAT_HELPERS.replace('__builtin_fmaf', 'fmaf')
    src += (_kernelgen.generate_cell_function(cm, prec)
            + _kernelgen.generate_split_cell_function(cm, prec) + WRAPPER)
    cpp = tmp_path / (tag + '.cpp')
    cpp.write_text(src)
    so = str(tmp_path / (tag + '.so'))
    r = subprocess.run(['g++', '-O1', '-fPIC', '-shared', '-std=c++17',
                        '-ffp-contract=off', '-I', CSRC, '-o', so, str(cpp)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    lib = ctypes.CDLL(so)
    rt = ctypes.c_double if prec == 64 else ctypes.c_float
    p = ctypes.POINTER(rt)
    for f in (lib.whole, lib.split):
        f.restype = None
        f.argtypes = [p, rt, rt, rt, rt, p, p]
    lib.qpart.restype = ctypes.c_int
    return lib, rt, (np.float64 if prec == 64 else np.float32)


def _compare(lib, rt, dtype, y0, nf, nq, seed):
    rng = np.random.default_rng(seed)
    p = ctypes.POINTER(rt)
    for trial in range(200):
        y = np.array(y0, dtype=dtype) * (1 + 0.05 * rng.standard_normal(
            len(y0))).astype(dtype)
        y[0] = dtype(rng.uniform(-90, 40))
        fv = (1 + 0.3 * rng.standard_normal(max(1, nf))).astype(dtype)
        t, dt = rt(rng.uniform(0, 500)), rt(0.01)
        pace = rt(float(rng.integers(0, 2)))
        idiff = rt(rng.uniform(-30, 30))
        ya, yb = y.copy(), y.copy()
        qa = np.zeros(max(1, nq), dtype=dtype)
        qb = np.zeros(max(1, nq), dtype=dtype)
        lib.whole(ya.ctypes.data_as(p), t, dt, pace, idiff,
                  fv.ctypes.data_as(p), qa.ctypes.data_as(p))
        lib.split(yb.ctypes.data_as(p), t, dt, pace, idiff,
                  fv.ctypes.data_as(p), qb.ctypes.data_as(p))
        ia = ya.view(np.uint64 if dtype == np.float64 else np.uint32)
        ib = yb.view(np.uint64 if dtype == np.float64 else np.uint32)
        assert np.array_equal(ia, ib), (trial, ya, yb)
        assert np.array_equal(qa[:nq], qb[:nq]), (trial, qa, qb)


@pytest.mark.parametrize('prec', [64, 32])
@pytest.mark.parametrize('rl', [False, True])
def test_split_equals_whole_lr1991(tmp_path, prec, rl):
    m, vm, rl_states, bound = _prepare_model(
        myokit.load_model('example'), rl, True)
    fields = [m.get('ina.gNa')]
    names = ['membrane.V', 'membrane.i_diff', 'ica.ICa', 'ina.m',
             'membrane.i_ion']
    quantities = [m.get(x) for x in names]
    cm = _kernelgen.CellModel(m, bound, rl_states, fields, quantities)
    dep, dstates = _kernelgen.diffusion_dependents(cm)
    # for LR1991 only the update of V needs the diffusion current
    assert [s.qname() for s in dstates] == ['membrane.V']
    lib, rt, dtype = _build(tmp_path, cm, prec, 'lr%d%d' % (prec, rl))
    # i_diff is the only logged quantity part B sets
    assert [lib.qpart(k) for k in range(len(names))] == [0, 1, 0, 0, 0]
    src = _kernelgen.generate_split_cell_function(cm, prec)
    assert '#define MK_NCARRY 2' in src      # i_ion, i_stim (V is in y)
    y0 = [s.initial_value(as_float=True) for s in cm.states]
    _compare(lib, rt, dtype, y0, 1, len(names), 7)


@pytest.mark.parametrize('prec', [64, 32])
def test_split_equals_whole_with_dependent_state(tmp_path, prec):
    m, vm, rl_states, bound = _prepare_model(
        myokit.parse_model(SIDE_MODEL), False, True)
    quantities = [m.get('c.s'), m.get('c.ion'), m.get('c.idiff')]
    cm = _kernelgen.CellModel(m, bound, rl_states, [], quantities)
    dep, dstates = _kernelgen.diffusion_dependents(cm)
    assert sorted(v.qname() for v in dep) == ['c.idiff', 'c.s']
    assert sorted(s.qname() for s in dstates) == ['c.u', 'c.v']
    lib, rt, dtype = _build(tmp_path, cm, prec, 'side%d' % prec)
    assert [lib.qpart(k) for k in range(3)] == [1, 0, 1]
    y0 = [s.initial_value(as_float=True) for s in cm.states]
    _compare(lib, rt, dtype, y0, 0, 3, 11)
