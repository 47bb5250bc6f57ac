"""
Discrete adjoint (vector-Jacobian product) of the generated cell step.

The reference has no adjoint; this builds one from the model's expression
tree, in the spirit of myokit's symbolic derivatives (``Expression.diff``,
``myokit/_expressions.py:243-336``) but in reverse mode. For one cell and one
step ``y' = F(y, idiff, fields)`` (the forward-Euler / Rush-Larsen update of
``_sim/openclsim.cl:358-364``), ``mk_cell_vjp`` returns, for a given
adjoint ``lam`` of ``y'``:

    mu      = (dF/dy)^T lam        (adjoint of the cell's own state)
    abar    = (dF/didiff)^T lam    (adjoint of the diffusion current)
    fbar[k] = (dF/dfield_k)^T lam

It recomputes the forward sweep, then walks the equations in reverse
dependency order. Each equation's expression tree is differentiated with the
ordinary rules of calculus. Contributions that flow through a branch of
``if``/``piecewise`` are added as ``cond ? contribution : 0``, so values of
the branch not taken (which may be inf or NaN) never reach the result.
Discontinuities are ignored, as in ``Expression.diff``.
"""
from . import _expressions as ex
from ._kernelgen import (DOUBLE_PRECISION, _cname, cell_writer,
                         forward_sweep, real_literal)


def _nonneg(e, cm, seen=None):
    """True if the value of ``e`` can be shown to be >= 0 (or NaN) from its
    form alone: non-negative numbers, exp, sqrt, and sums, products and
    quotients of those, following intermediate variables."""
    if isinstance(e, ex.Number):
        return e.eval() >= 0
    if isinstance(e, (ex.Exp, ex.Sqrt)):
        return True
    if isinstance(e, (ex.Plus, ex.Multiply, ex.Divide)):
        return _nonneg(e[0], cm, seen) and _nonneg(e[1], cm, seen)
    if isinstance(e, ex.Power) and isinstance(e[1], ex.Number):
        x = e[1].eval()
        return x == int(x) and int(x) % 2 == 0 or _nonneg(e[0], cm, seen)
    if isinstance(e, ex.Name):
        var = e.var()
        if var in cm.constants:
            return cm.constants[var] >= 0
        seen = set() if seen is None else seen
        if var in cm.order and not var.is_state() and var not in seen:
            seen.add(var)
            return _nonneg(var.rhs(), cm, seen)
    return False


def _may_cancel(e, cm, seen=None):
    """True if a multiplicative factor of ``e`` is a sum or difference
    that can cancel to zero (following intermediate variables)."""
    if seen is None:
        seen = set()
    if isinstance(e, ex.Minus):
        return True
    if isinstance(e, ex.Plus):
        return not (_nonneg(e[0], cm) and _nonneg(e[1], cm))
    if isinstance(e, ex.Multiply):
        return _may_cancel(e[0], cm, seen) or _may_cancel(e[1], cm, seen)
    if isinstance(e, (ex.Divide, ex.PrefixMinus, ex.PrefixPlus)):
        return _may_cancel(e[0], cm, seen)
    if isinstance(e, ex.Power) and isinstance(e[1], ex.Number) \
            and e[1].eval() > 0:
        return _may_cancel(e[0], cm, seen)
    if isinstance(e, ex.Name):
        var = e.var()
        if var in cm.order and not var.is_state() and var not in seen:
            seen.add(var)
            return _may_cancel(var.rhs(), cm, seen)
    return False


def _variable(e, cm):
    """True if ``e`` depends on anything but literal constants."""
    if isinstance(e, ex.Number):
        return False
    if isinstance(e, ex.Name):
        return e.var() not in cm.constants
    return any(_variable(x, cm) for x in e)


def needs_island(e, cm):
    """True if ``e`` holds a quotient N / D with a denominator factor that
    can cancel and a numerator that is not constant: the form of a
    removable singularity, such as 0.32 (V + 47.13) / (1 - exp(-0.1 (V +
    47.13))) or (exp(0.04 (V + 77)) - 1) / ((V + 77) exp(...)). Near it the
    quotient rule subtracts two large, nearly equal terms, and a
    single-precision VJP keeps few correct digits of the derivative."""
    if isinstance(e, ex.Divide) and _may_cancel(e[1], cm) \
            and _variable(e[0], cm):
        return True
    return any(needs_island(x, cm) for x in e)


class _Rev:
    """Reverse-mode code emitter for one cell."""

    def __init__(self, cm, precision):
        self.cm = cm
        self.p = precision
        self.w = cell_writer(cm, precision, False)
        if precision == DOUBLE_PRECISION:
            # the VJP's forward sweep needs no exactly rounded division
            self.w.rcp64 = 'mk_rcp64'
        self.bars = {}          # LhsExpression -> C name of its adjoint
        states = set(cm.states)
        for s in cm.states:
            self.bars[ex.Name(s)] = _cname(s, 'bx_')
            self.bars[ex.Derivative(ex.Name(s))] = _cname(s, 'bd_')
        for var in cm.order:
            if var not in states:
                self.bars[ex.Name(var)] = _cname(var, 'bv_')
        for var in cm.per_cell_constants:
            self.bars[ex.Name(var)] = _cname(var, 'bc_')
        for k, var in enumerate(cm.fields):
            self.bars[ex.Name(var)] = 'bf' + str(k)
        for var, what in cm.bound.items():
            if what == 'idiff':
                self.bars[ex.Name(var)] = 'b_idiff'
        self.out = []
        self._ntmp = 0
        self._inf = {}          # Variable -> may its value be infinite

    def lit(self, v):
        return real_literal(v, self.p)

    def v(self, e):
        """C code of the forward value of sub-expression ``e``."""
        return self.w.ex(e)

    def tmp(self, code):
        """Binds ``code`` to a new local and returns its name."""
        name = 't' + str(self._ntmp)
        self._ntmp += 1
        self.out.append('    const Real ' + name + ' = ' + code + ';')
        return name

    def may_inf(self, e):
        """False if the forward value of ``e`` is finite whenever the
        states, fields and inputs are: no exp, division, power, logarithm,
        tangent or remainder below it."""
        if isinstance(e, ex.Number):
            x = e.eval()
            return x != x or abs(x) == float('inf')
        if isinstance(e, ex.Derivative):
            return True
        if isinstance(e, ex.Name):
            var = e.var()
            if not hasattr(var, 'is_state') or var.is_state() or \
                    var in self.cm.fields or var in self.cm.bound:
                return False
            r = self._inf.get(var)
            if r is None:
                self._inf[var] = True       # provisional, for safety
                r = self._inf[var] = self.may_inf(var.rhs())
            return r
        if isinstance(e, (ex.Exp, ex.Divide, ex.Power, ex.Log, ex.Log10,
                          ex.Tan, ex.Quotient, ex.Remainder)):
            return True
        return any(self.may_inf(x) for x in e)

    def mul(self, g, v, e=None):
        """Adjoint ``g`` times forward value ``v`` (the value of ``e``). If
        that value may be infinite, the product is zero when ``g`` is zero
        (an overflowed forward value met by a zero adjoint gives 0, not
        0 * inf = NaN)."""
        if e is not None and not self.may_inf(e):
            return '(' + g + ' * ' + v + ')'
        return 'mk_gmul(' + g + ', ' + v + ')'

    def recip(self, b):
        """1 / b: single precision by the hardware reciprocal, double
        precision by the reciprocal with two Newton steps (mk_rcp64)."""
        if self.p == DOUBLE_PRECISION:
            return 'mk_rcp64(' + b + ')'
        return '__builtin_amdgcn_rcpf(' + b + ')'

    def div(self, a, b):
        if self.p == DOUBLE_PRECISION:
            return '(' + a + ' / ' + b + ')'
        return '(' + a + ' * __builtin_amdgcn_rcpf(' + b + '))'

    def gdiv(self, g, b):
        """Adjoint ``g`` divided by forward value ``b``, zero when ``g`` is
        zero."""
        return 'mk_gdiv(' + g + ', ' + b + ')'

    def acc(self, target, contrib, guard):
        if guard:
            self.out.append('    ' + target + ' += (' + ' && '.join(guard)
                            + ') ? (' + contrib + ') : (Real)0;')
        else:
            self.out.append('    ' + target + ' += ' + contrib + ';')

    def rev(self, e, g, guard=()):
        """Adds d(output)/d(leaves of e) given the adjoint ``g`` of e."""
        if isinstance(e, ex.Number):
            return
        if isinstance(e, (ex.Name, ex.Derivative)):
            b = self.bars.get(e)
            if b is not None:
                self.acc(b, g, guard)
            return
        if isinstance(e, ex.PrefixPlus):
            return self.rev(e[0], g, guard)
        if isinstance(e, ex.PrefixMinus):
            return self.rev(e[0], '(-' + g + ')', guard)
        if isinstance(e, ex.Plus):
            self.rev(e[0], g, guard)
            return self.rev(e[1], g, guard)
        if isinstance(e, ex.Minus):
            self.rev(e[0], g, guard)
            return self.rev(e[1], '(-' + g + ')', guard)
        if isinstance(e, ex.Multiply):
            a, b = e
            self.rev(a, self.mul(g, self.v(b), b), guard)
            return self.rev(b, self.mul(g, self.v(a), a), guard)
        if isinstance(e, ex.Divide):
            a, b = e
            # one reciprocal for both operands: d/da = 1/b,
            # d/db = -(a/b)/b
            rb = self.tmp(self.recip(self.v(b)))
            gt = self.tmp(g)
            self.rev(a, self.mul(gt, rb), guard)
            return self.rev(b, self.mul('(-' + gt + ')', '(' + self.v(e)
                                        + ' * ' + rb + ')'), guard)
        if isinstance(e, (ex.Quotient, ex.Floor, ex.Ceil)):
            return
        if isinstance(e, ex.Remainder):
            a, b = e
            self.rev(a, g, guard)
            fl = '__builtin_floor' if self.p == DOUBLE_PRECISION else 'floorf'
            return self.rev(b, self.mul('(-' + g + ')', fl + '('
                            + self.div(self.v(a), self.v(b)) + ')'), guard)
        if isinstance(e, ex.Power):
            a, b = e
            if isinstance(b, ex.Number):
                n = b.eval()
                if n == int(n) and abs(n - 1) <= 64:
                    pw = self.w.fn('powi') + '(' + self.v(a) + ', ' \
                        + str(int(n) - 1) + ')'
                else:
                    pw = self.w.fn('pow') + '(' + self.v(a) + ', ' \
                        + self.lit(n - 1) + ')'
                return self.rev(a, self.mul(g, '(' + self.lit(n) + ' * '
                                            + pw + ')'), guard)
            # a^b: d/da = b a^(b-1) = b (a^b) / a; d/db = a^b log(a)
            ve = self.v(e)
            self.rev(a, self.mul(g, '(' + self.v(b) + ' * '
                                 + self.div(ve, self.v(a)) + ')'), guard)
            return self.rev(b, self.mul(g, '(' + ve + ' * '
                                        + self.w.fn('log') + '('
                                        + self.v(a) + '))'), guard)
        if isinstance(e, ex.Exp):
            return self.rev(e[0], self.mul(g, self.v(e)), guard)
        if isinstance(e, ex.Log):
            if len(e) == 2:
                return self.rev(ex.Divide(ex.Log(e[0]), ex.Log(e[1])), g,
                                guard)
            return self.rev(e[0], self.gdiv(g, self.v(e[0])), guard)
        if isinstance(e, ex.Log10):
            return self.rev(e[0], self.gdiv(
                g, '(' + self.v(e[0]) + ' * ' + self.lit(2.302585092994046)
                + ')'), guard)
        if isinstance(e, ex.Sqrt):
            return self.rev(e[0], self.gdiv(
                g, '(' + self.lit(2) + ' * ' + self.v(e) + ')'), guard)
        if isinstance(e, ex.Sin):
            return self.rev(e[0], self.mul(g, self.w.fn('cos') + '('
                                           + self.v(e[0]) + ')'), guard)
        if isinstance(e, ex.Cos):
            return self.rev(e[0], self.mul('(-' + g + ')', self.w.fn('sin')
                                           + '(' + self.v(e[0]) + ')'), guard)
        if isinstance(e, ex.Tan):
            c = self.w.fn('cos') + '(' + self.v(e[0]) + ')'
            return self.rev(e[0], self.gdiv(g, '(' + c + ' * ' + c + ')'),
                            guard)
        if isinstance(e, (ex.ASin, ex.ACos)):
            a = self.v(e[0])
            s = self.mul(g, self.div(self.lit(1), self.w.fn('sqrt') + '('
                                     + self.lit(1) + ' - ' + a + ' * ' + a
                                     + ')'))
            if isinstance(e, ex.ACos):
                s = '(-' + s + ')'
            return self.rev(e[0], s, guard)
        if isinstance(e, ex.ATan):
            a = self.v(e[0])
            return self.rev(e[0], self.gdiv(
                g, '(' + self.lit(1) + ' + ' + a + ' * ' + a + ')'), guard)
        if isinstance(e, ex.Abs):
            a = self.v(e[0])
            return self.rev(e[0], '((' + a + ' >= ' + self.lit(0) + ') ? '
                            + g + ' : (-' + g + '))', guard)
        if isinstance(e, ex.If):
            c = self.v(e[0])
            self.rev(e[1], g, tuple(guard) + ('(' + c + ')',))
            return self.rev(e[2], g, tuple(guard) + ('(!' + c + ')',))
        if isinstance(e, ex.Piecewise):
            ops = list(e)
            prev = []
            for i in range(0, len(ops) - 1, 2):
                c = self.v(ops[i])
                self.rev(ops[i + 1], g, tuple(guard) + tuple(prev)
                         + ('(' + c + ')',))
                prev.append('(!' + c + ')')
            return self.rev(ops[-1], g, tuple(guard) + tuple(prev))
        if isinstance(e, ex.Condition):
            return
        raise ValueError('No adjoint rule for ' + type(e).__name__)


_VJP_EXP = """
/* exp in double precision for the VJP (no bit-exact counterpart is needed
 * here): x = k ln2 + r, |r| <= ln2 / 2, with the two-part ln2 of
 * Cody and Waite applied by fused multiply-adds, exp(r) by its Taylor
 * series to degree 11 (truncation below 2e-15 relative for |r| <= ln2/2),
 * and 2^k by ldexp, which gives inf and 0 beyond the range. */
__device__ __forceinline__ double mk_exp_vjp(double x)
{
    /* beyond these bounds the result is 0 or inf anyway (a NaN argument
     * gives 0: it can only come from a forward state that is NaN) */
    x = __builtin_fmin(__builtin_fmax(x, -760.0), 720.0);
    const double k = __builtin_rint(x * 0x1.71547652b82fep+0);
    const double r = __builtin_fma(-k, 0x1.abc9e3b39803fp-56,
                                   __builtin_fma(-k, 0x1.62e42fefa39efp-1,
                                                 x));
    double p = 0x1.ae64567f544e4p-26;                 /* 1/11! */
    p = __builtin_fma(p, r, 0x1.27e4fb7789f5cp-22);   /* 1/10! */
    p = __builtin_fma(p, r, 0x1.71de3a556c734p-19);   /* 1/9!  */
    p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-16);   /* 1/8!  */
    p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-13);   /* 1/7!  */
    p = __builtin_fma(p, r, 0x1.6c16c16c16c17p-10);   /* 1/6!  */
    p = __builtin_fma(p, r, 0x1.1111111111111p-7);    /* 1/5!  */
    p = __builtin_fma(p, r, 0x1.5555555555555p-5);    /* 1/4!  */
    p = __builtin_fma(p, r, 0x1.5555555555555p-3);    /* 1/3!  */
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return __builtin_ldexp(p, (int)k);
}
"""

_MIXED_WRAPPER = """
/* Single-precision state, double-precision VJP: the adjoint through an
 * upstroke amplifies the VJP's own rounding by orders of magnitude (the
 * quotient rule near removable singularities such as x / (1 - exp(-x))
 * cancels in single precision), so the products are formed in double and
 * only the results are rounded back. */
__device__ __forceinline__ void mk_cell_vjp(
    const Real* __restrict__ y, const Real t, const Real dt, const Real pace,
    const Real idiff, const Real* __restrict__ fv,
    const Real* __restrict__ lam, Real* __restrict__ mu,
    Real* __restrict__ abar, Real* __restrict__ fbar)
{
    double yd[MK_NS], ld[MK_NS], md[MK_NS], fd[MK_NFA], fbd[MK_NFA], ad;
#pragma unroll
    for (int i = 0; i < MK_NS; ++i) {
        yd[i] = (double)y[i];
        ld[i] = (double)lam[i];
    }
#pragma unroll
    for (int k = 0; k < MK_NFA; ++k) fd[k] = k < MK_NF ? (double)fv[k] : 0.0;
    mk_cell_vjp_f64(yd, (double)t, (double)dt, (double)pace, (double)idiff,
                    fd, ld, md, &ad, fbd);
#pragma unroll
    for (int i = 0; i < MK_NS; ++i) mu[i] = (Real)md[i];
    *abar = (Real)ad;
#pragma unroll
    for (int k = 0; k < MK_NF; ++k) fbar[k] = (Real)fbd[k];
}
"""


_VJP_RCP64 = """
/* 1/b in double precision for the VJP: the hardware reciprocal refined by
 * two Newton steps, r <- r + r (1 - b r); for b = 0, inf or NaN the
 * refinement would give NaN, so the plain reciprocal is returned. */
__device__ __forceinline__ double mk_rcp64(double b)
{
    const double r0 = __builtin_amdgcn_rcp(b);
    double r = __builtin_fma(__builtin_fma(-b, r0, 1.0), r0, r0);
    r = __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
    return __builtin_isfinite(r) ? r : r0;
}
"""


_ISLAND_RCP = """
/* 1/b for the double-precision islands: the hardware reciprocal and one
 * Newton step (error squared; ample for 1e-11); for b = 0, inf or NaN the
 * plain reciprocal. */
__device__ __forceinline__ double mk_rcp_isl(double b)
{
    const double r0 = __builtin_amdgcn_rcp(b);
    const double r = __builtin_fma(__builtin_fma(-b, r0, 1.0), r0, r0);
    return __builtin_isfinite(r) ? r : r0;
}
"""


_VJP_ONEM_EXPF = """
/* 1 - exp(-a) in single precision, for the Rush-Larsen partial
 * d y' / d inf. Formed as 1 - E it loses log2(1/a) bits to cancellation when
 * dt << tau (a = dt / tau near 0), and the adjoint amplifies that error
 * through every upstroke. Below |a| = 0.5 the Taylor series of
 * (1 - exp(-a)) / a to degree 7 (truncation under 1e-8 relative) is used,
 * above it 1 - E with E = exp(-a), which is then exact to a few ulp. */
__device__ __forceinline__ float mk_onem_expf(const float a, const float E)
{
    float p = -0x1.a01a02p-16f;
    p = __builtin_fmaf(p, a, 0x1.a01a02p-13f);
    p = __builtin_fmaf(p, a, -0x1.6c16c2p-10f);
    p = __builtin_fmaf(p, a, 0x1.111112p-7f);
    p = __builtin_fmaf(p, a, -0x1.555556p-5f);
    p = __builtin_fmaf(p, a, 0x1.555556p-3f);
    p = __builtin_fmaf(p, a, -0.5f);
    p = __builtin_fmaf(p, a, 1.0f);
    return __builtin_fabsf(a) < 0.5f ? a * p : 1.0f - E;
}
"""


def generate_vjp_function(cm, precision, compute=None, islands=True):
    """HIP source of ``mk_cell_vjp`` for a CellModel. ``compute`` is the
    precision of the VJP's arithmetic (default: ``precision``); double
    arithmetic for a single-precision kernel gives a double-precision
    ``mk_cell_vjp_f64`` and a wrapper that converts on entry and exit.

    A single-precision VJP evaluates the reverse rule of each equation that
    holds a removable singularity (``needs_island``) in double precision,
    and rounds only the sums it adds to other adjoints (unless ``islands``
    is false)."""
    import re
    if compute is None or compute == precision:
        src = _vjp_source(cm, precision, islands)
        if precision == DOUBLE_PRECISION:
            return _VJP_EXP + _VJP_RCP64 + re.sub(r'\bmk_exp\(',
                                                'mk_exp_vjp(', src)
        head = _VJP_LMUL + _VJP_ONEM_EXPF
        if 'mk_rcp_isl' in src or 'mk_exp_isl' in src or '_f64(' in src:
            head = _ISLAND_EXP + _ISLAND_RCP + _ISLAND_GUARDS + head
        return head + src
    if not (precision != DOUBLE_PRECISION and compute == DOUBLE_PRECISION):
        raise ValueError('Only a double-precision VJP for a single-precision'
                         ' kernel can be mixed.')
    return _mixed_source(cm)


#: Single-precision multiply with 0 * x = 0 for every x, inf and NaN
#: included (v_mul_legacy_f32): the adjoint guard in one instruction
_VJP_LMUL = """
/* The AMDGPU intrinsic behind v_mul_legacy_f32, declared by its LLVM name:
 * the compiler schedules it like any VALU instruction and inserts the wait
 * states its operands need (a read of a transcendental's result, say). */
__device__ float mk_fmul_legacy(float, float)
    __asm("llvm.amdgcn.fmul.legacy");

__device__ __forceinline__ float mk_lmul(const float g, const float v)
{
    return mk_fmul_legacy(g, v);
}
"""


_ISLAND_EXP = """
/* exp in double precision for the double-precision islands of a
 * single-precision VJP, where 1e-11 relative is ample: x = k ln2 + r,
 * exp(r) = exp(r / 16)^16 with exp(r / 16) by its Taylor series to degree
 * 5 (truncation 1.5e-13, 2.3e-12 after the four squarings), 2^k by
 * ldexp. Three constants instead of eleven. */
__device__ __forceinline__ double mk_exp_isl(double x)
{
    x = __builtin_fmin(__builtin_fmax(x, -760.0), 720.0);
    const double k = __builtin_rint(x * 0x1.71547652b82fep+0);
    const double s = __builtin_fma(-k, 0x1.62e42fefa39efp-5, x * 0.0625);
    double p = __builtin_fma(s, 0x1.1111111111111p-7, 0x1.5555555555555p-5);
    p = __builtin_fma(p, s, 0x1.5555555555555p-3);
    p = __builtin_fma(p, s, 0.5);
    p = __builtin_fma(p, s, 1.0);
    p = __builtin_fma(p, s, 1.0);
    p = p * p;
    p = p * p;
    p = p * p;
    p = p * p;
    return __builtin_ldexp(p, (int)k);
}
"""


_ISLAND_GUARDS = """
__device__ __forceinline__ double mk_gmul_f64(const double g, const double v)
{
    return g == 0.0 ? 0.0 : g * v;
}

__device__ __forceinline__ double mk_gdiv_f64(const double g, const double b)
{
    return g == 0.0 ? 0.0 : g / b;
}
"""


def _island(r, e, k, tab):
    """Partial derivatives of ``e`` with respect to its leaves, computed in
    double precision right after the forward sweep (the reverse rule of
    ``e`` seeded with 1), where few single-precision values are live, and
    rounded once. Returns the lines that compute them and a map from each
    leaf's adjoint to the name of its partial derivative; the reverse
    sweep then adds ``adjoint(e) * partial`` in single precision."""
    import re
    r64 = _Rev(r.cm, DOUBLE_PRECISION)
    r64.w.names = {key: '((double)' + v + ')'
                   for key, v in r64.w.names.items()}
    r64._ntmp = r._ntmp
    r64._inf = r._inf
    for key in list(r64.bars):
        r64.bars[key] = 'isl_' + r.bars[key]
    # each distinct exp once (inside the branches of a piecewise the
    # compiler would not merge repeated evaluations), and exps of the same
    # rate, exp(a x + m2) = exp(a x + m1) exp(m2 - m1), from one evaluation
    import math
    exps = []
    rates = {}

    def find(x):
        for y in x:
            find(y)
        if isinstance(x, ex.Exp) and x.code() not in r64.w.subst:
            name = 'ie' + str(k) + '_' + str(len(exps))
            aff = r64.w.affine(x[0])
            base = rates.get(aff[:2]) if aff is not None else None
            if base is not None and abs(aff[2] - base[1]) <= 20:
                code = '(' + base[0] + ' * ' + real_literal(
                    math.exp(aff[2] - base[1]), DOUBLE_PRECISION) + ')'
            else:
                code = r64.w.ex(x)
                if aff is not None:
                    rates[aff[:2]] = (name, aff[2])
            exps.append('const double ' + name + ' = ' + code + ';')
            r64.w.subst[x.code()] = name
    find(e)
    r64.out.extend(tab + line for line in exps)
    r64.rev(e, '1.0')
    r._ntmp = r64._ntmp
    body = '\n'.join(r64.out)
    body = re.sub(r'\bReal\b', 'double', body)
    body = re.sub(r'\bmk_gmul\(', 'mk_gmul_f64(', body)
    body = re.sub(r'\bmk_gdiv\(', 'mk_gdiv_f64(', body)
    body = re.sub(r'\bmk_exp\(', 'mk_exp_isl(', body)
    body = re.sub(r'\bmk_rcp64\(', 'mk_rcp_isl(', body)
    used = sorted(set(re.findall(r'\bisl_\w+', body)))
    jac = {u[4:]: 'J' + str(k) + '_' + u[4:] for u in used}
    lines = [tab + 'Real ' + jac[u[4:]] + ';' for u in used]
    win = _island_window(r, e)
    if win is not None:
        lines.append(tab + 'if (' + win + ') {')
    else:
        lines.append(tab + '{')
    lines += [tab + tab + 'double ' + u + ' = 0.0;' for u in used]
    lines += [tab + line for line in body.split('\n') if line]
    lines += [tab + tab + jac[u[4:]] + ' = (Real)' + u + ';' for u in used]
    if win is not None:
        # Away from the singular point the single-precision rule is as
        # accurate as the rest of the VJP
        r32 = _Rev(r.cm, r.p)
        r32.w.set_exp_members(r.w.exp_members)
        r32._ntmp = r._ntmp
        r32._inf = r._inf
        for key in list(r32.bars):
            r32.bars[key] = 'isf_' + r.bars[key]
        r32.rev(e, '(Real)1')
        r._ntmp = r32._ntmp
        body32 = '\n'.join(r32.out)
        lines.append(tab + '} else {')
        lines += [tab + tab + 'Real isf_' + u[4:] + ' = (Real)0;'
                  for u in used]
        lines += [tab + line for line in body32.split('\n') if line]
        lines += [tab + tab + jac[u[4:]] + ' = isf_' + u[4:] + ';'
                  for u in used]
    lines.append(tab + '}')
    return lines, jac


# Half-width, in units of the exponent's argument, of the window around a
# removable singularity of the form N / (1 - exp(a x + m)) in which the
# partial derivatives are computed in double precision. The single-precision
# quotient rule loses ~6.5e-5 / u^2 relative at u mV from ina.m's point (a =
# 0.1 per mV): 3e-7 at the window's edge (|a x + m| = 1.5, u = 15 mV).
ISLAND_WINDOW = 1.5

# Half-width, in units of the variable, of the window around the root of a
# cancelling factor a x + m (ik.xi's V + 77 below its numerator
# exp(0.04 u) - 1): outside it the single-precision quotient rule's two
# terms, each ~ 1/u, cancel to ~0.02 with a relative error of about
# 3e-6 / u, under 1e-6 at 5 mV; its share of mu is at the 1e-8 level there
# (tests/test_adjoint_codegen.py). The resting potential (-84 mV) and the
# plateau lie outside, so most cells skip the island.
ISLAND_WINDOW_LINEAR = 5.0


def _island_window(r, e):
    """C condition (single precision) under which the island of ``e`` must
    be computed in double precision, or None if it always must: every
    cancelling factor of every denominator in ``e`` has the form 1 -
    exp(a x + m) or exp(a x + m) - 1, the condition then being
    |a x + m| < ISLAND_WINDOW, or the form a x + m, the condition then
    being |a x + m| < ISLAND_WINDOW_LINEAR |a|, with x one variable and a,
    m literals. Quotients with any other cancelling factor keep the island
    everywhere."""
    conds = []

    def factors(d):
        if isinstance(d, ex.Multiply):
            return factors(d[0]) + factors(d[1])
        return [d]

    def expm1_arg(f):
        if isinstance(f, ex.Minus):
            a, b = f
            if isinstance(a, ex.Number) and a.eval() == 1 and \
                    isinstance(b, ex.Exp):
                return b[0]
            if isinstance(b, ex.Number) and b.eval() == 1 and \
                    isinstance(a, ex.Exp):
                return a[0]
        return None

    def walk(x):
        if isinstance(x, ex.Divide) and _may_cancel(x[1], r.cm):
            for f in factors(x[1]):
                if not _may_cancel(f, r.cm):
                    continue
                arg = expm1_arg(f)
                aff = r.w.affine(arg) if arg is not None else None
                if aff is not None:
                    width = ISLAND_WINDOW
                else:
                    aff = r.w.affine(f)
                    if aff is None or aff[1] == 0:
                        return False
                    width = ISLAND_WINDOW_LINEAR * abs(aff[1])
                conds.append('__builtin_fabsf(__builtin_fmaf(' + aff[0]
                             + ', ' + real_literal(aff[1], r.p) + ', '
                             + real_literal(aff[2], r.p) + ')) < '
                             + real_literal(width, r.p))
        return all(walk(y) for y in x)
    if r.p == DOUBLE_PRECISION or not walk(e) or not conds:
        return None
    return ' || '.join(sorted(set(conds)))


def _mixed_source(cm):
    """Double-precision ``mk_cell_vjp_f64`` and the single-precision
    ``mk_cell_vjp`` wrapper around it."""
    import re
    src = _vjp_source(cm, DOUBLE_PRECISION)
    src = re.sub(r'\bReal\b', 'double', src)
    for name in ('mk_cell_vjp', 'mk_gmul', 'mk_gdiv'):
        src = re.sub(r'\b' + name + r'\b', name + '_f64', src)
    src = re.sub(r'\bmk_exp\(', 'mk_exp_vjp(', src)
    return _VJP_EXP + _VJP_RCP64 + src + _MIXED_WRAPPER


def _vjp_source(cm, precision, islands=True):
    """HIP source of ``mk_cell_vjp`` computing in ``precision``, with
    double-precision islands (``_island``) in single precision."""
    r = _Rev(cm, precision)
    w = r.w
    tab = '    '
    if precision == DOUBLE_PRECISION:
        gmul = 'g == (Real)0 ? (Real)0 : g * v'
        gdiv = 'g == (Real)0 ? (Real)0 : ' + r.div('g', 'b')
    else:
        # The legacy multiply gives 0 when either factor is 0, whatever the
        # other (inf and NaN included): the guard in one instruction
        gmul = 'mk_lmul(g, v)'
        gdiv = 'mk_lmul(g, __builtin_amdgcn_rcpf(b))'
    lines = [
        '__device__ __forceinline__ Real mk_gmul(const Real g, const Real v)',
        '{',
        tab + 'return ' + gmul + ';',
        '}',
        '',
        '__device__ __forceinline__ Real mk_gdiv(const Real g, const Real b)',
        '{',
        tab + 'return ' + gdiv + ';',
        '}',
        '',
        '__device__ __forceinline__ void mk_cell_vjp(',
        tab + 'const Real* __restrict__ y, const Real t, const Real dt,'
        ' const Real pace, const Real idiff,',
        tab + 'const Real* __restrict__ fv, const Real* __restrict__ lam,'
        ' Real* __restrict__ mu,',
        tab + 'Real* __restrict__ abar, Real* __restrict__ fbar)',
        '{',
        tab + '(void)t; (void)dt; (void)pace; (void)idiff; (void)fv;',
    ]
    fam = []
    if precision != DOUBLE_PRECISION:
        # the forward kernel's exponential families (one exp per family)
        from ._kernelgen import exp_families
        fam, members = exp_families(cm, w)
        w.set_exp_members(members)
    sweep = forward_sweep(cm, w, tab)
    ns = len(cm.states)
    lines.extend(sweep[:ns] + [tab + x for x in fam] + sweep[ns:])
    islands_jac = {}
    if islands and precision != DOUBLE_PRECISION:
        for var in cm.order:
            if not (var.is_state() and var in cm.rl_states) and \
                    needs_island(var.rhs(), cm):
                pre, jac = _island(r, var.rhs(), len(islands_jac), tab)
                lines.extend(pre)
                islands_jac[var] = jac
    for b in sorted(set(r.bars.values())):
        lines.append(tab + 'Real ' + b + ' = (Real)0;')
    # Outputs: the state update of every state. The direct terms of
    # mu = dy'/dy^T lam (lam for forward Euler, lam E for Rush-Larsen) are
    # kept out of the accumulators: the reverse sweep adds dozens of
    # contributions of order dt to each state's adjoint, and summed into an
    # accumulator that already holds lam, each would be rounded to lam's
    # precision. In single precision that is several ulp of lam per step,
    # the same at every step at rest, and so a drift of the adjoint.
    expf = w.fn('exp')
    direct = []
    for i, s in enumerate(cm.states):
        x = _cname(s, 'x_')
        if s in cm.rl_states:
            inf, tau = cm.rl_states[s]
            vi, vt = w.ex(ex.Name(inf)), w.ex(ex.Name(tau))
            E = 'E_' + str(i)
            if precision == DOUBLE_PRECISION:
                lines.append(tab + 'const Real ' + E + ' = ' + expf + '('
                             + r.div('-dt', vt) + ');')
                D = '((Real)1 - ' + E + ')'
            else:
                # E as the forward kernel computes it (generate_cell_function:
                # 2^((-dt log2(e)) X) with X = 1 / tau), A = dt X for 1 - E
                trhs = tau.rhs()
                if isinstance(trhs, ex.Divide) and isinstance(
                        trhs[0], ex.Number) and trhs[0].eval() == 1:
                    rate = w.ex(trhs[1])
                else:
                    rate = '__builtin_amdgcn_rcpf(' + vt + ')'
                A, D = 'A_' + str(i), 'D_' + str(i)
                lines.append(tab + 'const Real ' + A + ' = dt * ' + rate
                             + ';')
                lines.append(tab + 'const Real ' + E + ' = __builtin_amdgcn_'
                             'exp2f((-dt * 0x1.715476p+0f) * ' + rate + ');')
                lines.append(tab + 'const Real ' + D + ' = mk_onem_expf('
                             + A + ', ' + E + ');')
            # y' = inf - (inf - x) E: dy'/dx = E, dy'/dinf = 1 - E,
            # dy'/dtau = -(inf - x) E dt / tau^2. The direct term lam E is
            # added to mu at the end (see below).
            direct.append('lam[' + str(i) + '] * ' + E)
            lines.append(tab + r.bars[ex.Name(inf)] + ' += lam[' + str(i)
                         + '] * ' + D + ';')
            lines.append(tab + r.bars[ex.Name(tau)] + ' += -lam[' + str(i)
                         + '] * (' + vi + ' - ' + x + ') * ' + E + ' * '
                         + r.div('dt', '(' + vt + ' * ' + vt + ')') + ';')
        else:
            direct.append('lam[' + str(i) + ']')
            lines.append(tab + _cname(s, 'bd_') + ' += dt * lam[' + str(i)
                         + '];')
    # Reverse dependency order
    for var in reversed(cm.order):
        if var.is_state():
            if var in cm.rl_states:
                continue
            target = r.bars[ex.Derivative(ex.Name(var))]
        else:
            target = r.bars[ex.Name(var)]
        if var in islands_jac:
            for bar, J in sorted(islands_jac[var].items()):
                r.out.append(tab + bar + ' += ' + r.mul(target, J) + ';')
        else:
            r.rev(var.rhs(), target)
    for var in reversed(cm.per_cell_constants):
        r.rev(var.rhs(), r.bars[ex.Name(var)])
    lines.extend(r.out)
    for i, s in enumerate(cm.states):
        lines.append(tab + 'mu[' + str(i) + '] = ' + direct[i] + ' + '
                     + _cname(s, 'bx_') + ';')
    lines.append(tab + '*abar = ' + ('b_idiff' if 'b_idiff' in
                                     r.bars.values() else '(Real)0') + ';')
    for k in range(len(cm.fields)):
        lines.append(tab + 'fbar[' + str(k) + '] = bf' + str(k) + ';')
    lines.append('}')
    return '\n'.join(lines) + '\n'
