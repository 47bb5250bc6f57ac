"""
CPU checks of the generated cell VJP (``mk_cell_vjp``, myokit_amd/_adjointgen):
the generated HIP function is compiled as C with gcc and compared with central
finite differences of the oracle's one-cell update (``oracle_cell``, the
forward-Euler / Rush-Larsen step of ``_sim/openclsim.cl:298-367``):

    mu = (dF/dy)^T lam,  abar = (dF/didiff)^T lam,  fbar = (dF/dfield)^T lam.

The single-precision variant runs with host stand-ins for the two AMDGCN
builtins it uses (exp2 and reciprocal), and is checked against the double
variant, including at rest where LR1991's ``ina.a`` switch overflows exp in
single precision.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import myokit_amd as myokit
from myokit_amd import _kernelgen
from myokit_amd._adjointgen import _VJP_LMUL, generate_vjp_function
from myokit_amd.simulation import _prepare_model

from oracle import cgen

CSRC = os.path.join(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__))), 'myokit_amd', 'csrc')

PRELUDE = '''
#include <math.h>
#include <stdbool.h>
#include "mk_math.h"
#define __device__ static
#define __forceinline__ inline
#define __restrict__ restrict
#define __builtin_amdgcn_exp2f(x) exp2f(x)
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))
#define __builtin_amdgcn_rcp(x) (1.0 / (x))
typedef %s Real;
'''

LMUL_HOST = '''
static inline float mk_lmul(float g, float v)
{
    return (g == 0.0f || v == 0.0f) ? 0.0f : g * v;
}
'''

WRAPPER = '''
void vjp(const Real* y, Real t, Real dt, Real pace, Real idiff,
         const Real* fv, const Real* lam, Real* mu, Real* abar, Real* fbar)
{
    mk_cell_vjp(y, t, dt, pace, idiff, fv, lam, mu, abar, fbar);
}
'''


def _gcc(src, path):
    c = str(path) + '.c'
    with open(c, 'w') as f:
        f.write(src)
    so = str(path) + '.so'
    r = subprocess.run(['gcc', '-O1', '-fPIC', '-shared', '-std=gnu11',
                        '-ffp-contract=off', '-I', CSRC, '-o', so, c, '-lm'],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return ctypes.CDLL(so)


class Cell:
    """A prepared model with its oracle step and compiled VJPs."""

    def __init__(self, tmp_path, model, rl, fields, diffusion=True):
        m, vm, rl_states, bound = _prepare_model(model, rl, diffusion)
        self.fields = [m.get(f) for f in fields]
        self.cm = _kernelgen.CellModel(m, bound, rl_states, self.fields, [])
        self.ns = len(self.cm.states)
        self.nf = len(self.fields)
        self.y0 = np.array([s.initial_value(as_float=True)
                            for s in self.cm.states])
        src = cgen.model_source(m, bound, rl_states, self.fields, [])
        self.orc = _gcc(src, tmp_path / 'orc')
        self.orc.oracle_cell.restype = None
        d = ctypes.POINTER(ctypes.c_double)
        self.orc.oracle_cell.argtypes = [
            d, ctypes.c_double, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, d, d, ctypes.c_int]
        self.vjps = {}
        defs = '#define MK_NS %d\n#define MK_NF %d\n#define MK_NFA %d\n' % (
            self.ns, self.nf, max(1, self.nf))
        for prec, ct, rt in ((64, 'double', ctypes.c_double),
                             (32, 'float', ctypes.c_float),
                             ('plain', 'float', ctypes.c_float),
                             ('mixed', 'float', ctypes.c_float)):
            src = PRELUDE % ct + defs
            if prec != 64:
                src += _kernelgen._FLOAT_HELPERS
            if prec == 'mixed':
                src += generate_vjp_function(self.cm, 32, 64) + WRAPPER
            elif prec == 'plain':
                src += generate_vjp_function(self.cm, 32, islands=False) \
                    + WRAPPER
            else:
                src += generate_vjp_function(self.cm, prec) + WRAPPER
            # host stand-in for the legacy multiply (0 * x = 0 for every x)
            src = src.replace(_VJP_LMUL, LMUL_HOST)
            lib = _gcc(src, tmp_path / ('vjp%s' % prec))
            p = ctypes.POINTER(rt)
            lib.vjp.restype = None
            lib.vjp.argtypes = [p, rt, rt, rt, rt, p, p, p, p, p]
            self.vjps[prec] = (lib, rt)

    def step(self, y, t, dt, pace, idiff, fv):
        y = np.array(y, dtype=np.float64)
        fv = np.ascontiguousarray(fv, dtype=np.float64)
        q = np.zeros(1)
        d = ctypes.POINTER(ctypes.c_double)
        self.orc.oracle_cell(y.ctypes.data_as(d), t, dt, pace, idiff,
                             fv.ctypes.data_as(d), q.ctypes.data_as(d), 0)
        return y

    def vjp(self, prec, y, t, dt, pace, idiff, fv, lam):
        lib, rt = self.vjps[prec]
        npt = np.float64 if prec == 64 else np.float32
        p = ctypes.POINTER(rt)
        y = np.ascontiguousarray(y, dtype=npt)
        fv = np.ascontiguousarray(fv if len(fv) else [0], dtype=npt)
        lam = np.ascontiguousarray(lam, dtype=npt)
        mu = np.zeros(self.ns, dtype=npt)
        ab = np.zeros(1, dtype=npt)
        fb = np.zeros(max(1, self.nf), dtype=npt)
        lib.vjp(y.ctypes.data_as(p), t, dt, pace, idiff, fv.ctypes.data_as(p),
                lam.ctypes.data_as(p), mu.ctypes.data_as(p),
                ab.ctypes.data_as(p), fb.ctypes.data_as(p))
        return mu.astype(float), float(ab[0]), fb[:self.nf].astype(float)

    def fd(self, y, t, dt, pace, idiff, fv, lam, rel=1e-4):
        """Fourth-order central differences of lam . F around (y, idiff,
        fv): (8 (phi(h) - phi(-h)) - (phi(2h) - phi(-2h))) / 12h."""
        def phi(yy, ii, ff):
            return float(np.dot(lam, self.step(yy, t, dt, pace, ii, ff)))

        def d(f, h):
            return (8 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12 * h)

        def shifted(v, k, e):
            v = np.array(v, dtype=float)
            v[k] += e
            return v
        hy = np.array([rel * max(abs(y[k]), 1e-2) for k in range(self.ns)])
        hi = rel * max(abs(idiff), 1.0)
        # Fields enter through small currents: a larger step keeps their
        # effect on V (~1e-11 for 1e-4) far above V's rounding (~1e-14)
        hf = np.array([1e-2 * max(abs(fv[k]), 1e-2) for k in range(self.nf)])
        mu = np.array([d(lambda e: phi(shifted(y, k, e), idiff, fv), hy[k])
                       for k in range(self.ns)])
        ab = d(lambda e: phi(y, idiff + e, fv), hi)
        fb = np.array([d(lambda e: phi(y, idiff, shifted(fv, k, e)), hf[k])
                       for k in range(self.nf)])
        # Rounding floor of each difference: phi carries an error of about
        # 2 eps sum |lam_i F_i|, which the formula scales by 1.5 / h
        err = 2 * 2.3e-16 * float(np.dot(np.abs(lam), np.abs(
            self.step(y, t, dt, pace, idiff, fv))))
        self.noise = (1.5 * err / hy, 1.5 * err / hi, 1.5 * err / hf)
        return mu, ab, fb


def _states(cell):
    """Resting state, a depolarised state and a mid-upstroke state."""
    y0 = cell.y0
    iv = [s.qname() for s in cell.cm.states].index('membrane.V')
    ys = [y0.copy()]
    y = y0.copy()
    y[iv] = 20.0
    ys.append(y)
    y = y0.copy()
    y[iv] = -40.5
    ys.append(y)
    return ys


def _close(a, b, scale, rtol):
    return np.all(np.abs(a - b) <= rtol * (np.abs(b) + scale))


@pytest.mark.parametrize('rl', [False, True])
def test_vjp_matches_finite_differences_double(tmp_path, rl):
    cell = Cell(tmp_path, myokit.load_model('example'), rl,
                ['ina.gNa', 'ica.gCa'])
    rng = np.random.default_rng(3)
    fv = np.array([16.0 * 1.1, 0.09 * 0.9])
    for y in _states(cell):
        for pace, idiff in ((0.0, 0.0), (1.0, -3.0)):
            lam = rng.normal(size=cell.ns)
            mu, ab, fb = cell.vjp(64, y, 1.0, 0.01, pace, idiff, fv, lam)
            fmu, fab, ffb = cell.fd(y, 1.0, 0.01, pace, idiff, fv, lam)
            nmu, nab, nfb = cell.noise
            assert np.all(np.abs(mu - fmu) <= 1e-6 * np.abs(fmu) + nmu), \
                (mu, fmu, nmu)
            assert abs(ab - fab) <= 1e-6 * abs(fab) + nab
            assert np.all(np.abs(fb - ffb) <= 1e-6 * np.abs(ffb) + nfb), \
                (fb, ffb, nfb)


def test_vjp_single_precision_no_overflow_nan(tmp_path):
    # At rest (V = -84 mV), exp(-(V + 40) / 0.24) = exp(183) overflows in
    # single precision inside ina.a = 1 - 1 / (1 + exp(...)): the adjoint
    # reaching that exp is exactly zero and must stay zero, not 0 * inf
    cell = Cell(tmp_path, myokit.load_model('example'), True,
                ['ina.gNa', 'ica.gCa'])
    rng = np.random.default_rng(5)
    fv = np.array([16.0, 0.09])
    for y in _states(cell):
        lam = rng.normal(size=cell.ns)
        m32, a32, f32 = cell.vjp(32, y, 1.0, 0.01, 0.0, 0.5, fv, lam)
        m64, a64, f64 = cell.vjp(64, y, 1.0, 0.01, 0.0, 0.5, fv, lam)
        assert np.all(np.isfinite(m32)) and np.isfinite(a32)
        assert np.all(np.isfinite(f32))
        assert _close(m32, m64, np.max(np.abs(m64)), 2e-3)
        assert a32 == pytest.approx(a64, rel=2e-3)
        assert _close(f32, f64, np.max(np.abs(f64)), 2e-3)


def test_vjp_guards_untaken_branches(tmp_path):
    # The untaken branch divides by zero at x = 0 (inf); its contribution
    # is guarded, so the adjoint is that of the taken branch only. sqrt,
    # log, power and trigonometric rules are exercised on the other state.
    src = '''[[model]]
c.x = 0
c.z = 0.5
[c]
t = 0 bind time
dot(x) = if(x == 0, 2 * z, 1 / x) + piecewise(z < 0, 1 / (x * z), z > 1, 3, z^2)
dot(z) = sqrt(z) + log(1 + z) * sin(z) + cos(x) * atan(z) + z^2.5 + exp(-z) / (1 + z)
'''
    m = myokit.parse_model(src)
    cell = Cell(tmp_path, m, False, [], diffusion=False)
    y = np.array([0.0, 0.5])
    lam = np.array([0.7, -1.3])
    for prec in (64, 32):
        mu, ab, _ = cell.vjp(prec, y, 0.0, 0.1, 0.0, 0.0, [], lam)
        assert np.all(np.isfinite(mu))
        # F = y + dt f(y); df_x/dz = 2 + 2 z at x = 0
        dz = lam[1] * 0.1 * (0.5 / np.sqrt(0.5) + np.log(1.5) * np.cos(0.5)
                             + np.sin(0.5) / 1.5 + 1.0 / (1 + 0.25)
                             + 2.5 * 0.5 ** 1.5
                             + (-np.exp(-0.5) * 1.5 - np.exp(-0.5)) / 2.25)
        exp_z = lam[1] + lam[0] * 0.1 * (2 + 2 * 0.5) + dz
        assert mu[1] == pytest.approx(exp_z, rel=1e-5 if prec == 32 else 1e-12)
        # d/dx: the taken branch 2 z has none; cos(x) atan(z) gives -sin(0)
        assert mu[0] == pytest.approx(lam[0], rel=1e-6)


def test_mixed_precision_vjp_matches_double(tmp_path):
    # Single-precision inputs and outputs, double-precision arithmetic: the
    # result is the double VJP of the rounded inputs, rounded once
    cell = Cell(tmp_path, myokit.load_model('example'), True,
                ['ina.gNa', 'ica.gCa'])
    rng = np.random.default_rng(8)
    fv = np.array([16.0, 0.09])
    for y in _states(cell):
        # near the removable singularity of ina.m's alpha (V = -47.13)
        for v in (y[0], -47.1299, -47.13):
            yy = y.copy()
            yy[0] = v
            lam = rng.normal(size=cell.ns) * 100
            y32 = yy.astype(np.float32).astype(float)
            l32 = lam.astype(np.float32).astype(float)
            f32 = fv.astype(np.float32).astype(float)
            mm, am, fm = cell.vjp('mixed', yy, 1.0, 0.01, 0.0, 0.5, fv, lam)
            m64, a64, f64 = cell.vjp(64, y32, np.float32(1.0),
                                     float(np.float32(0.01)), 0.0, 0.5, f32,
                                     l32)
            np.testing.assert_allclose(mm, np.float32(m64), rtol=1e-6,
                                       atol=1e-6 * np.max(np.abs(m64)))
            assert am == pytest.approx(float(np.float32(a64)), rel=1e-6)
            np.testing.assert_allclose(fm, np.float32(f64), rtol=1e-6,
                                       atol=1e-6 * np.max(np.abs(f64)))


def test_vjp_exp_against_libm(tmp_path):
    # The double-precision exp of the mixed VJP (not the bit-exact forward
    # one): relative error against the C library over the double range,
    # and the limits
    from myokit_amd._adjointgen import _VJP_EXP
    src = '#include <math.h>\n#define __device__ static\n' \
          '#define __forceinline__ inline\n' + _VJP_EXP + '''
void run(const double* x, double* y, long n)
{
    for (long i = 0; i < n; ++i) y[i] = mk_exp_vjp(x[i]);
}
'''
    lib = _gcc(src, tmp_path / 'expvjp')
    d = ctypes.POINTER(ctypes.c_double)
    lib.run.argtypes = [d, d, ctypes.c_long]
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-745, 709, 200000),
                        rng.uniform(-1e-3, 1e-3, 100000),
                        [0.0, -800.0, 800.0, np.inf, -np.inf, np.nan]])
    y = np.zeros_like(x)
    lib.run(x.ctypes.data_as(d), y.ctypes.data_as(d), len(x))
    with np.errstate(over='ignore'):
        ref = np.exp(x)
    ok = np.isfinite(ref) & (ref >= 2.0 ** -1022)
    assert np.max(np.abs(y[ok] - ref[ok]) / ref[ok]) < 2e-14
    assert y[-6] == 1.0 and y[-5] == 0.0 and np.isinf(y[-4])
    assert np.isinf(y[-3]) and y[-2] == 0.0


def test_single_precision_rl_partials_at_rest(tmp_path):
    # At rest the slow gates have dt / tau ~ 1e-4: d y' / dV contains
    # inf'(V) (1 - E) with E = exp(-dt / tau) ~ 1, which formed as 1 - E in
    # single precision keeps only a few correct digits. The VJP forms it as
    # -expm1(-dt / tau): every d gate' / dV agrees with the double VJP of
    # the same inputs to 1e-5 relative.
    cell = Cell(tmp_path, myokit.load_model('example'), True,
                ['ina.gNa', 'ica.gCa'])
    fv = np.array([16.0, 0.09])
    y = cell.y0.astype(np.float32).astype(float)
    iv = [s.qname() for s in cell.cm.states].index('membrane.V')
    for k, s in enumerate(cell.cm.states):
        if s not in cell.cm.rl_states:
            continue
        lam = np.zeros(cell.ns)
        lam[k] = 1.0
        m32 = cell.vjp(32, y, 1.0, 0.01, 0.0, 0.0, fv, lam)[0]
        m64 = cell.vjp(64, y, float(np.float32(1.0)),
                       float(np.float32(0.01)), 0.0, 0.0,
                       fv.astype(np.float32).astype(float), lam)[0]
        assert m32[iv] == pytest.approx(m64[iv], rel=1e-5), s.qname()


def test_onem_expf_against_libm(tmp_path):
    # 1 - exp(-a) of the single-precision VJP against double expm1, over
    # the series range, across its switch point and beyond
    from myokit_amd._adjointgen import _VJP_ONEM_EXPF
    src = '#include <math.h>\n#define __device__ static\n' \
          '#define __forceinline__ inline\n' \
          '#define __builtin_fmaf fmaf\n#define __builtin_fabsf fabsf\n' \
          + _VJP_ONEM_EXPF + '''
void run(const float* a, float* y, long n)
{
    for (long i = 0; i < n; ++i) y[i] = mk_onem_expf(a[i], expf(-a[i]));
}
'''
    lib = _gcc(src, tmp_path / 'onem')
    f = ctypes.POINTER(ctypes.c_float)
    lib.run.argtypes = [f, f, ctypes.c_long]
    rng = np.random.default_rng(2)
    a = np.concatenate([10.0 ** rng.uniform(-9, 1.5, 100000),
                        -(10.0 ** rng.uniform(-9, 1, 20000)),
                        [0.0, 0.49999997, 0.5, 0.50000006, np.inf]])
    a = a.astype(np.float32)
    y = np.zeros_like(a)
    lib.run(a.ctypes.data_as(f), y.ctypes.data_as(f), len(a))
    ref = -np.expm1(-a.astype(float))
    ok = np.isfinite(ref) & (ref != 0)
    assert np.max(np.abs(y[ok] - ref[ok]) / np.abs(ref[ok])) < 4e-7
    assert y[-5] == 0.0 and y[-1] == 1.0


def test_island_exp_against_libm(tmp_path):
    # The double-precision exp of the islands of the single-precision VJP:
    # 1e-11 relative over the double range, and the limits
    from myokit_amd._adjointgen import _ISLAND_EXP
    src = '#include <math.h>\n#define __device__ static\n' \
          '#define __forceinline__ inline\n' + _ISLAND_EXP + \
          '''
void run(const double* x, double* y, long n)
{
    for (long i = 0; i < n; ++i) y[i] = mk_exp_isl(x[i]);
}
'''
    lib = _gcc(src, tmp_path / 'expisl')
    d = ctypes.POINTER(ctypes.c_double)
    lib.run.argtypes = [d, d, ctypes.c_long]
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.uniform(-745, 709, 200000),
                        rng.uniform(-1e-3, 1e-3, 100000),
                        [0.0, -800.0, 800.0, np.inf, -np.inf]])
    y = np.zeros_like(x)
    lib.run(x.ctypes.data_as(d), y.ctypes.data_as(d), len(x))
    with np.errstate(over='ignore'):
        ref = np.exp(x)
    ok = np.isfinite(ref) & (ref >= 2.0 ** -1022)
    assert np.max(np.abs(y[ok] - ref[ok]) / ref[ok]) < 1e-11
    assert y[-5] == 1.0 and y[-4] == 0.0 and np.isinf(y[-3])
    assert np.isinf(y[-2]) and y[-1] == 0.0


def test_legacy_multiply_is_compiler_visible(tmp_path):
    # The adjoint guards (0 * x = 0 for every x) are v_mul_legacy_f32 through
    # the llvm.amdgcn.fmul.legacy intrinsic, not inline assembly: the
    # compiler then inserts the wait states their operands need (on gfx950
    # a VALU read of a transcendental's result needs one). The reverse
    # kernels of a single-precision adjoint are compiled here (hiprtc needs
    # no GPU): they hold no inline assembly, and the guards reach the code
    # object as legacy multiplies (or legacy multiply-adds).
    from myokit_amd import _native
    m = myokit.load_model('example')
    s = myokit.SimulationOpenCL(m, None, ncells=(64, 32), precision=32,
                                rl=True)
    s.set_field('ina.gNa', np.ones((32, 64)))
    cm = _kernelgen.CellModel(s._model, s._bound_variables, s._rl_states,
                              [s._model.get('ina.gNa')],
                              [s._model.get('membrane.V')])
    src = _kernelgen.generate_kernel(cm, 32, 1, adjoint=True)
    assert 'asm(' not in src.replace('__asm("llvm.amdgcn.fmul.legacy")', '')
    opts = _kernelgen.compile_options(32) + ' --save-temps'
    prog = ctypes.c_void_p()
    log = ctypes.create_string_buffer(1 << 16)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        _native.check(_native.lib().mkh_program_build(
            src.encode(), opts.encode(), None, -1, ctypes.byref(prog), log,
            1 << 16), 'mkh_program_build')
    finally:
        os.chdir(cwd)
    _native.lib().mkh_program_free(prog)
    asm = [f for f in os.listdir(tmp_path) if f.endswith('.s')]
    assert asm
    text = open(os.path.join(tmp_path, asm[0])).read()
    lines = [x.strip() for x in text.splitlines()
             if x.strip() and not x.strip().startswith((';', '.'))]
    n = sum(1 for x in lines if x.startswith(
        ('v_mul_legacy_f32', 'v_mad_legacy_f32', 'v_fma_legacy_f32')))
    assert n >= 3 * 50      # three reverse kernels, dozens of guards each


def test_reverse_sheet_tile_height_and_lds_budget():
    # The generator writes the reverse sheet's cells per thread as MK_ACPT
    # and sizes its workgroups per CU from the tile's LDS: mu (8 states) of
    # a 64 x 4 cpt tile and abar with a one-cell ring, in 160 KiB per CU
    m = myokit.load_model('example')
    s = myokit.SimulationOpenCL(m, None, ncells=(64, 32), precision=32,
                                rl=True)
    cm = _kernelgen.CellModel(s._model, s._bound_variables, s._rl_states,
                              [], [s._model.get('membrane.V')])
    for cpt, groups in ((8, 2), (4, 4), (2, 4)):
        lds = 8 * 4 * cpt * 64 * 4 + (4 * cpt + 2) * 66 * 4 + 68
        assert groups == min(4, 160 * 1024 // lds)
        assert _kernelgen.adjoint_sheet_groups(cm, 32, cpt) == groups
        src = _kernelgen.generate_kernel(cm, 32, 1, adjoint=True,
                                         adjoint_cpt=cpt)
        assert '#define MK_ADJ_SHEET 1' in src
        assert '#define MK_ACPT %d' % cpt in src
        assert '#define MK_ADJ_SWAVES %d' % groups in src
    src = _kernelgen.generate_kernel(cm, 32, 1, adjoint=True)
    assert '#define MK_ACPT %d' % _kernelgen.adjoint_sheet_cpt(32) in src
    with pytest.raises(ValueError, match='adjoint_cpt'):
        _kernelgen.generate_kernel(cm, 32, 1, adjoint=True, adjoint_cpt=3)
    with pytest.raises(ValueError, match='adjoint_cpt'):
        s.set_tuning(adjoint_cpt=16)
    s.set_tuning(adjoint_cpt=8)
    assert s.tuning()['adjoint_cpt'] == 8


def test_island_window_edges(tmp_path):
    # ina.m's alpha, 0.32 (V + 47.13) / (1 - exp(-0.1 (V + 47.13))), is
    # differentiated in double precision only within 15 mV of its singular
    # point (|-0.1 (V + 47.13)| < 1.5) and by the single-precision rule
    # outside; ik.xi, whose cancelling factor is V + 77, within 5 mV of -77
    # mV. On both sides of the windows' edges and far from the
    # singular point, d m' / dV of the single-precision VJP agrees with the
    # double one to 1e-5 relative. Within 0.3 mV of the point the single-
    # precision forward values themselves (1 - exp(-0.1 u) for small u) carry
    # ~1e-5: 1e-4 there. (At the point itself, within ~1e-4 mV, the quotient
    # rule cancels to second order in any precision, as alpha does in the
    # forward step: the model's own conditioning there, not tested.)
    from myokit_amd import _adjointgen
    model = myokit.load_model('example')
    cell = Cell(tmp_path, model, True, ['ina.gNa'])
    src = generate_vjp_function(cell.cm, 32)
    assert src.count('if (__builtin_fabsf(__builtin_fmaf(') == 2
    assert _adjointgen.ISLAND_WINDOW == 1.5
    assert _adjointgen.ISLAND_WINDOW_LINEAR == 5.0
    fv = np.array([16.0])
    names = [s.qname() for s in cell.cm.states]
    iv, im = names.index('membrane.V'), names.index('ina.m')
    lam = np.zeros(cell.ns)
    lam[im] = 1.0
    for dv in (-40, -15.2, -15.0, -14.8, -5, -0.3, -0.01, 0.01, 0.3, 5,
               14.8, 15.0, 15.2, 40):
        y = cell.y0.copy()
        y[iv] = -47.13 + dv
        y = y.astype(np.float32).astype(float)
        m32 = cell.vjp(32, y, 1.0, 0.01, 0.0, 0.0, fv, lam)[0]
        m64 = cell.vjp(64, y, float(np.float32(1.0)), float(np.float32(0.01)),
                       0.0, 0.0, fv, lam)[0]
        tol = 1e-4 if abs(dv) <= 0.3 else 1e-5
        assert m32[iv] == pytest.approx(m64[iv], rel=tol), dv
    # ik.xi's window: the partial d V' / dV with ik.x open (IK at full
    # weight), across the edges at -77 +- 5 mV, at rest (-84 mV) and on the
    # plateau
    ix = names.index('ik.x')
    lam = np.zeros(cell.ns)
    lam[iv] = 1.0
    for dv in (-40, -7, -5.2, -5.0, -4.8, -1, -0.01, 0.01, 1, 4.8, 5.0, 5.2,
               7, 80):
        y = cell.y0.copy()
        y[iv] = -77.0 + dv
        y[ix] = 1.0
        y = y.astype(np.float32).astype(float)
        m32 = cell.vjp(32, y, 1.0, 0.01, 0.0, 0.0, fv, lam)[0]
        m64 = cell.vjp(64, y, float(np.float32(1.0)), float(np.float32(0.01)),
                       0.0, 0.0, fv, lam)[0]
        assert m32[iv] == pytest.approx(m64[iv], rel=1e-5), dv
