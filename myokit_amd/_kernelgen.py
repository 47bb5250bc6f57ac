"""
HIP code generation for the tissue engine.

``HipExpressionWriter`` turns model expressions into HIP C++ (the role of
``myokit.formats.opencl.OpenCLExpressionWriter``, ``formats/opencl/
_ewriter.py:12-111``, on top of the C semantics of ``formats/ansic/
_ewriter.py:12-121``: ``pow``, ``floor``-based quotient and remainder,
ternary ``if``/``piecewise``, ``log(a)/log(b)``). ``generate_kernel`` writes
one device function ``mk_cell`` that evaluates the model and advances one
cell by one step (the body of ``cell_step``, ``_sim/openclsim.cl:298-367``),
then appends the fixed driver kernels of ``csrc/mk_drivers.hip``.

Differences from the reference's kernel text, by design:

* constants (literal and calculated, unless a field depends on them) are
  evaluated once on the host in double precision and written as exact
  hexadecimal literals (the reference writes ``#define`` expressions that the
  device evaluates, ``openclsim.cl:173-186``);
* double precision uses the portable functions of ``csrc/mk_math.h`` so the
  device matches the CPU oracle bit for bit; single precision uses the
  device's float functions (``__expf`` etc. when ``native_maths`` is set, the
  analogue of OpenCL's ``native_*``);
* integer powers with a literal exponent ``|n| <= 64`` use binary powering
  (``mk_powi``) instead of a general ``pow``;
* state is read from and written to SoA registers, not AoS memory.
"""
import math
import os

import numpy as np

from . import _expressions as ex

_CSRC = os.path.join(os.path.dirname(__file__), 'csrc')

SINGLE_PRECISION = 32
DOUBLE_PRECISION = 64


def _read(name):
    with open(os.path.join(_CSRC, name), 'r') as f:
        return f.read()


def real_literal(value, precision):
    """Exact C literal for ``value`` in the given precision."""
    value = float(value)
    if precision == SINGLE_PRECISION:
        value = float(np.float32(value))
    if math.isnan(value):
        return '(Real)__builtin_nan("")'
    if math.isinf(value):
        return '(Real)__builtin_inf()' if value > 0 \
            else '(-(Real)__builtin_inf())'
    s = value.hex()
    if precision == SINGLE_PRECISION:
        s += 'f'
    if value < 0 or s.startswith('-'):
        return '(' + s + ')'
    return s


# Function names per precision: (double, float, float-native)
_FUNCS = {
    'sqrt': ('__builtin_sqrt', 'sqrtf', 'sqrtf'),
    'exp': ('mk_exp', 'mk_expf', 'mk_expf_native'),
    'log': ('mk_log', 'logf', '__logf'),
    'log10': ('mk_log10', 'log10f', '__log10f'),
    'sin': ('mk_sin', 'sinf', '__sinf'),
    'cos': ('mk_cos', 'cosf', '__cosf'),
    'tan': ('mk_tan', 'tanf', '__tanf'),
    'asin': ('mk_asin', 'asinf', 'asinf'),
    'acos': ('mk_acos', 'acosf', 'acosf'),
    'atan': ('mk_atan', 'atanf', 'atanf'),
    'floor': ('__builtin_floor', 'floorf', 'floorf'),
    'ceil': ('__builtin_ceil', 'ceilf', 'ceilf'),
    'abs': ('__builtin_fabs', 'fabsf', 'fabsf'),
    'pow': ('mk_pow', 'powf', '__powf'),
    'powi': ('mk_powi', 'mk_powif', 'mk_powif'),
}


class HipExpressionWriter:
    """Writes expressions as HIP C++ source text.

    ``names`` maps LhsExpressions to the C text that stands for them; a
    constant without an entry is inlined as its (host-evaluated) value.
    """

    def __init__(self, precision=SINGLE_PRECISION, native_maths=False):
        self.precision = precision
        self.native = bool(native_maths)
        self.names = {}
        self.constants = {}
        # sub-expression code -> C name of a local holding its value
        self.subst = {}
        # exp() code -> (C name of a power B^n, factor exp(m)), single
        # precision exponential families
        self.exp_members = {}

    def set_exp_members(self, members):
        """Writes each exponential in ``members`` (code -> (C name of B^n,
        exp(m)), from :func:`exp_families`) as that power times the factor,
        folding the factor into a literal coefficient where the
        exponential is multiplied or divided by one."""
        self.exp_members = dict(members)

    def _member(self, e, coef=1.0):
        p, f = self.exp_members[e.code()]
        f *= coef
        if f == 1.0:
            return p
        return '(' + p + ' * ' + real_literal(f, self.precision) + ')'

    def fn(self, name):
        d, f, fn = _FUNCS[name]
        if self.precision == DOUBLE_PRECISION:
            return d
        return fn if self.native else f

    def ex(self, e):
        if self.subst:
            name = self.subst.get(e.code())
            if name is not None:
                return name
        m = getattr(self, '_ex_' + type(e).__name__)
        return m(e)

    # Atoms
    def _ex_Number(self, e):
        return real_literal(e.eval(), self.precision)

    def _ex_Name(self, e):
        try:
            return self.names[e]
        except KeyError:
            pass
        var = e.var()
        if var in self.constants:
            return real_literal(self.constants[var], self.precision)
        raise KeyError('No C name for <' + e.code() + '>')

    def _ex_Derivative(self, e):
        return self.names[e]

    # Operators
    def _ex_PrefixPlus(self, e):
        return '(+' + self.ex(e[0]) + ')'

    def _ex_PrefixMinus(self, e):
        return '(-' + self.ex(e[0]) + ')'

    def _infix(self, e, op):
        return '(' + self.ex(e[0]) + ' ' + op + ' ' + self.ex(e[1]) + ')'

    def _ex_Plus(self, e):
        return self._infix(e, '+')

    def _ex_Minus(self, e):
        return self._infix(e, '-')

    def _ex_Multiply(self, e):
        if self.exp_members:
            for a, b in ((e[0], e[1]), (e[1], e[0])):
                if isinstance(a, ex.Exp) and isinstance(b, ex.Number) \
                        and a.code() in self.exp_members:
                    return self._member(a, b.eval())
        return self._infix(e, '*')

    # Double precision divides with IEEE '/' (bit-exact with the oracle).
    # Set to a function name (e.g. 'mk_rcp64') to write a * f(b) instead,
    # where exact rounding is not required (the adjoint's forward sweep).
    rcp64 = None

    def _ex_Divide(self, e):
        if self.precision == DOUBLE_PRECISION:
            if self.rcp64 is None:
                return self._infix(e, '/')
            a, b = e
            if isinstance(b, ex.Number):
                return ('(' + self.ex(a) + ' * '
                        + real_literal(1.0 / b.eval(), self.precision) + ')')
            return ('(' + self.ex(a) + ' * ' + self.rcp64 + '(' + self.ex(b)
                    + '))')
        # Single precision: a * rcp(b) (v_rcp_f32, within the 2.5 ulp that
        # OpenCL allows for single-precision division); a literal divisor
        # becomes a multiplication by its reciprocal.
        a, b = e
        if isinstance(b, ex.Number) and isinstance(a, ex.Exp) \
                and a.code() in self.exp_members:
            return self._member(a, 1.0 / b.eval())
        if isinstance(b, ex.Number):
            return ('(' + self.ex(a) + ' * '
                    + real_literal(1.0 / b.eval(), self.precision) + ')')
        if isinstance(a, ex.Number) and a.eval() == 1.0:
            return '__builtin_amdgcn_rcpf(' + self.ex(b) + ')'
        return ('(' + self.ex(a) + ' * __builtin_amdgcn_rcpf(' + self.ex(b)
                + '))')

    def _ex_Quotient(self, e):
        # Rounds towards minus infinity, as myokit.Quotient
        return self.fn('floor') + '(' + self._infix(e, '/') + ')'

    def _ex_Remainder(self, e):
        a, b = self.ex(e[0]), self.ex(e[1])
        return ('(' + a + ' - ' + b + ' * ' + self.fn('floor') + '(' + a
                + ' / ' + b + '))')

    def _ex_Power(self, e):
        a, b = e
        if isinstance(b, ex.Number):
            v = b.eval()
            if v == int(v) and abs(v) <= 64:
                return (self.fn('powi') + '(' + self.ex(a) + ', '
                        + str(int(v)) + ')')
        return self.fn('pow') + '(' + self.ex(a) + ', ' + self.ex(b) + ')'

    # Functions
    def _func(self, name, e):
        return self.fn(name) + '(' + ', '.join(self.ex(x) for x in e) + ')'

    def _ex_Sqrt(self, e):
        return self._func('sqrt', e)

    def _ex_Exp(self, e):
        if self.exp_members and e.code() in self.exp_members:
            return self._member(e)
        if self.precision == SINGLE_PRECISION:
            aff = self.affine(e[0])
            if aff is not None:
                # exp(k*x + m) = 2^(x * k*log2(e) + m*log2(e)): one fma and
                # one v_exp_f32. The rounding error of the fused argument is
                # no larger than that of evaluating k*x + m in single
                # precision, which the reference's float kernel also has.
                x, k, m = aff
                l2e = 1.4426950408889634
                return ('__builtin_amdgcn_exp2f(__builtin_fmaf(' + x + ', '
                        + real_literal(k * l2e, self.precision) + ', '
                        + real_literal(m * l2e, self.precision) + '))')
        return self._func('exp', e)

    def affine(self, e):
        """If ``e`` is ``k * x + m`` for one name ``x`` (not a host
        constant) and literal ``k``, ``m``, returns ``(c_name_of_x, k, m)``
        with k, m in double precision; else None."""
        r = self._affine(e)
        if r is None or r[0] is None or r[1] == 0:
            return None
        return r

    def _affine(self, e):
        # Returns (name_text or None, k, m)
        if isinstance(e, ex.Number):
            return (None, 0.0, e.eval())
        if isinstance(e, ex.Name):
            var = e.var()
            if e not in self.names and var in self.constants:
                return (None, 0.0, float(self.constants[var]))
            return (self.ex(e), 1.0, 0.0)
        if isinstance(e, ex.PrefixPlus):
            return self._affine(e[0])
        if isinstance(e, ex.PrefixMinus):
            a = self._affine(e[0])
            return None if a is None else (a[0], -a[1], -a[2])
        if isinstance(e, (ex.Plus, ex.Minus)):
            a, b = self._affine(e[0]), self._affine(e[1])
            if a is None or b is None:
                return None
            if a[0] is not None and b[0] is not None and a[0] != b[0]:
                return None
            sg = 1.0 if isinstance(e, ex.Plus) else -1.0
            return (a[0] or b[0], a[1] + sg * b[1], a[2] + sg * b[2])
        if isinstance(e, (ex.Multiply, ex.Divide)):
            a, b = self._affine(e[0]), self._affine(e[1])
            if a is None or b is None:
                return None
            if isinstance(e, ex.Divide):
                if b[0] is not None or b[2] == 0:
                    return None
                return (a[0], a[1] / b[2], a[2] / b[2])
            if a[0] is None:
                return (b[0], a[2] * b[1], a[2] * b[2])
            if b[0] is None:
                return (a[0], a[1] * b[2], a[2] * b[2])
            return None
        return None

    def _ex_Log(self, e):
        if len(e) == 1:
            return self._func('log', e)
        f = self.fn('log')
        return ('(' + f + '(' + self.ex(e[0]) + ') / ' + f + '('
                + self.ex(e[1]) + '))')

    def _ex_Log10(self, e):
        return self._func('log10', e)

    def _ex_Sin(self, e):
        return self._func('sin', e)

    def _ex_Cos(self, e):
        return self._func('cos', e)

    def _ex_Tan(self, e):
        return self._func('tan', e)

    def _ex_ASin(self, e):
        return self._func('asin', e)

    def _ex_ACos(self, e):
        return self._func('acos', e)

    def _ex_ATan(self, e):
        return self._func('atan', e)

    def _ex_Floor(self, e):
        return self._func('floor', e)

    def _ex_Ceil(self, e):
        return self._func('ceil', e)

    def _ex_Abs(self, e):
        return self._func('abs', e)

    def _ex_If(self, e):
        return ('(' + self.ex(e[0]) + ' ? ' + self.ex(e[1]) + ' : '
                + self.ex(e[2]) + ')')

    def _ex_Piecewise(self, e):
        ops = list(e)
        out = []
        for i in range(0, len(ops) - 1, 2):
            out.append('(' + self.ex(ops[i]) + ' ? ' + self.ex(ops[i + 1])
                       + ' : ')
        out.append(self.ex(ops[-1]))
        out.append(')' * ((len(ops) - 1) // 2))
        return ''.join(out)

    # Conditions
    def _ex_Not(self, e):
        return '(!' + self.ex(e[0]) + ')'

    def _ex_Equal(self, e):
        return self._infix(e, '==')

    def _ex_NotEqual(self, e):
        return self._infix(e, '!=')

    def _ex_More(self, e):
        return self._infix(e, '>')

    def _ex_Less(self, e):
        return self._infix(e, '<')

    def _ex_MoreEqual(self, e):
        return self._infix(e, '>=')

    def _ex_LessEqual(self, e):
        return self._infix(e, '<=')

    def _ex_And(self, e):
        return self._infix(e, '&&')

    def _ex_Or(self, e):
        return self._infix(e, '||')


class CellModel:
    """The per-cell computation extracted from a prepared model.

    ``model``      cloned model (RL-converted if needed, bindings prepared,
                   unique names created)
    ``bound``      {variable: 'time' | 'pace' | 'idiff'}
    ``rl_states``  {state: (inf_variable, tau_variable)}
    ``fields``     ordered variables replaced by per-cell values
    ``quantities`` ordered per-cell logged variables (states, intermediaries,
                   or the variable bound to the diffusion current)
    """

    def __init__(self, model, bound, rl_states, fields, quantities):
        self.model = model
        self.bound = dict(bound)
        self.rl_states = dict(rl_states)
        self.fields = list(fields)
        self.quantities = list(quantities)
        self.states = list(model.states())
        for q in self.quantities:
            if not (q.is_state() or q.is_intermediary() or q in self.bound):
                raise ValueError('Cannot log <' + q.qname() + '>.')

        # Constants: host-evaluated unless they depend on a field
        fielded = set(self.fields)
        self.per_cell_constants = []
        self.constants = {}
        subst = {}
        for var in model.constants_order():
            if var in fielded:
                continue
            deps = [r.var() for r in var.rhs().references()]
            if any(d in fielded for d in deps) or any(
                    d in set(self.per_cell_constants) for d in deps):
                self.per_cell_constants.append(var)
                continue
            value = var.rhs()._eval(subst)
            subst[ex.Name(var)] = value
            self.constants[var] = value
        self.order = model.solvable_order()

    def literal_constants(self):
        return dict(self.constants)


def _cname(var, prefix='v_'):
    return prefix + var.uname()


def cell_writer(cm, precision, native_maths=False):
    """A HipExpressionWriter with the C names of one cell's variables."""
    w = HipExpressionWriter(precision, native_maths)
    w.constants = cm.constants
    bound_names = {'time': 't', 'pace': 'pace', 'idiff': 'idiff'}
    for var, what in cm.bound.items():
        w.names[ex.Name(var)] = bound_names[what]
    for i, s in enumerate(cm.states):
        w.names[ex.Name(s)] = _cname(s, 'x_')
        w.names[ex.Derivative(ex.Name(s))] = _cname(s, 'd_')
    for k, var in enumerate(cm.fields):
        w.names[ex.Name(var)] = 'fv[' + str(k) + ']'
    for var in cm.per_cell_constants:
        w.names[ex.Name(var)] = _cname(var, 'c_')
    for var in cm.order:
        if not var.is_state():
            w.names[ex.Name(var)] = _cname(var)
    return w


def forward_sweep(cm, w, tab='    '):
    """C statements evaluating every equation of one cell (states read from
    ``y``), in dependency order."""
    lines = []
    for i, s in enumerate(cm.states):
        lines.append(tab + 'const Real ' + _cname(s, 'x_') + ' = y['
                     + str(i) + '];')
    for var in cm.per_cell_constants:
        lines.append(tab + 'const Real ' + _cname(var, 'c_') + ' = '
                     + w.ex(var.rhs()) + ';')
    for var in cm.order:
        if var.is_state():
            if var in cm.rl_states:
                continue
            lines.append(tab + 'const Real ' + _cname(var, 'd_') + ' = '
                         + w.ex(var.rhs()) + ';')
        else:
            lines.append(tab + 'const Real ' + _cname(var) + ' = '
                         + w.ex(var.rhs()) + ';')
    return lines


def exp_families(cm, w, max_power=24, v_range=200.0):
    """Single precision: groups the exponentials exp(k V + m) of the
    membrane potential V whose rates are integer multiples k = n k0 (n = 1
    .. max_power) of a common rate k0, such as exp(-0.01 V), exp(-0.02 V),
    exp(-0.04 V) and exp(-0.1 (V + 47.13)) in LR1991. A family evaluates
    B = exp(k0 V) once (one v_exp_f32) and each member as B^n exp(m), the
    powers by multiplications: one transcendental per family instead of one
    per member, at a rounding error of about n ulp (below that of the
    argument k V + m itself when |k V + m| is large).

    A member is admitted only where nothing can leave the single-precision
    range for |V| <= ``v_range``: n |k0| v_range <= 60 and |m| <= 40.

    Returns ``(lines, members)``: the statements computing B and its
    powers (to follow the state loads) and a map from each member's
    expression code to ``(C name of B^n, exp(m))``, for
    :meth:`HipExpressionWriter.set_exp_members`."""
    vm = cm.model.label('membrane_potential')
    if vm is None or vm not in cm.states:
        return [], {}
    xname = w.names[ex.Name(vm)]
    found = {}
    # exponentials multiplied or divided by a literal, which can absorb
    # the factor exp(m)
    scalable = set()

    def walk(e):
        for x in e:
            walk(x)
        if isinstance(e, ex.Exp):
            aff = w.affine(e[0])
            if aff is not None and aff[0] == xname:
                found.setdefault(e.code(), (aff[1], aff[2]))
        elif isinstance(e, ex.Multiply):
            for a, b in ((e[0], e[1]), (e[1], e[0])):
                if isinstance(a, ex.Exp) and isinstance(b, ex.Number):
                    scalable.add(a.code())
        elif isinstance(e, ex.Divide):
            if isinstance(e[0], ex.Exp) and isinstance(e[1], ex.Number):
                scalable.add(e[0].code())
    for var in cm.order:
        walk(var.rhs())
    found = {c: km for c, km in found.items() if abs(km[1]) <= 40}
    remaining = set(k for k, m in found.values())
    lines, subst = [], {}
    nfam = 0

    def multiple(k, k0):
        r = k / k0
        n = int(round(r))
        if n < 1 or n > max_power or abs(r - n) > 1e-9 * n:
            return None
        if n * abs(k0) * v_range > 60:
            return None
        return n
    while remaining:
        best = None
        for k0 in sorted(remaining, key=abs):
            mem = {k: multiple(k, k0) for k in remaining}
            mem = {k: n for k, n in mem.items() if n is not None}
            count = sum(1 for k, m in found.values() if k in mem)
            if best is None or count > best[0]:
                best = (count, k0, mem)
        count, k0, mem = best
        remaining -= set(mem) | {k0}
        if count < 2:
            continue
        base = 'xe' + str(nfam)
        # powers by an addition chain: each new power is the product of two
        # known ones where possible, else by the binary method
        have = {1: base}
        plines = []
        for n in sorted(set(mem.values())):
            if n in have:
                continue
            pair = None
            for a in sorted(have, reverse=True):
                if n - a in have:
                    pair = (a, n - a)
                    break
            if pair is None:
                steps = []
                q = n
                while q not in have:
                    steps.append(q)
                    q = q // 2 if q % 2 == 0 else q - 1
                for q in reversed(steps):
                    a = q // 2 if q % 2 == 0 else q - 1
                    have[q] = base + '_' + str(q)
                    plines.append('const Real ' + have[q] + ' = ' + have[a]
                                  + ' * ' + have[q - a] + ';')
                continue
            have[n] = base + '_' + str(n)
            plines.append('const Real ' + have[n] + ' = ' + have[pair[0]]
                          + ' * ' + have[pair[1]] + ';')
        # Issue cost: a member evaluated directly is one fma and one
        # v_exp_f32, priced 4 + 16 cycles; in the family it is a share of
        # the powers (4 each) and, if m != 0 and no literal factor absorbs
        # exp(m), one multiply (4). Keep the family if it gains. (The SQ
        # counters show transcendentals at half rate on gfx950, 8 cycles;
        # at that price LR1991's two-member family exp(0.05 V) only breaks
        # even, and every other family still gains.)
        members = [(c, k, m) for c, (k, m) in found.items() if k in mem]
        saved = 20 * (len(members) - 1)
        added = 4 * len(plines) + 4 * sum(
            1 for c, k, m in members if m != 0 and c not in scalable)
        if added >= saved:
            continue
        nfam += 1
        lines.append('const Real ' + base + ' = __builtin_amdgcn_exp2f('
                     + xname + ' * '
                     + real_literal(k0 * 1.4426950408889634, SINGLE_PRECISION)
                     + ');')
        lines.extend(plines)
        for code, k, m in sorted(members):
            subst[code] = (have[mem[k]], math.exp(m))
    return lines, subst


def generate_cell_function(cm, precision, native_maths=False):
    """Returns the HIP source of ``mk_cell<LOG>`` for a CellModel."""
    w = cell_writer(cm, precision, native_maths)
    lines = []
    tab = '    '

    lines.append('template <bool LOG>')
    lines.append('__device__ __forceinline__ void mk_cell(')
    lines.append(tab + 'Real* __restrict__ y, const Real t, const Real dt,'
                 ' const Real pace,')
    lines.append(tab + 'const Real idiff, const Real* __restrict__ fv,'
                 ' Real* __restrict__ q)')
    lines.append('{')
    lines.append(tab + '(void)t; (void)dt; (void)pace; (void)idiff;'
                 ' (void)fv; (void)q;')
    fam = []
    if precision == SINGLE_PRECISION:
        fam, members = exp_families(cm, w)
        w.set_exp_members(members)
    sweep = forward_sweep(cm, w, tab)
    ns = len(cm.states)
    lines.extend(sweep[:ns] + [tab + x for x in fam] + sweep[ns:])
    if cm.quantities:
        lines.append(tab + 'if (LOG) {')
        for k, var in enumerate(cm.quantities):
            lines.append(2 * tab + 'q[' + str(k) + '] = '
                         + w.ex(ex.Name(var)) + ';')
        lines.append(tab + '}')
    for i, s in enumerate(cm.states):
        lines.append(tab + _state_update(cm, w, precision, i, s))
    lines.append('}')
    return '\n'.join(lines) + '\n'


def _state_update(cm, w, precision, i, s):
    """The C statement that advances state ``s`` (index ``i``) by one step:
    y[i] = x + dt * dot(x), or the Rush-Larsen update."""
    x = _cname(s, 'x_')
    rl = cm.rl_states
    if s not in rl:
        return 'y[' + str(i) + '] = ' + x + ' + dt * ' + _cname(s, 'd_') + ';'
    inf, tau = rl[s]
    inf, tau = w.ex(ex.Name(inf)), w.ex(ex.Name(tau))
    # RL update x = inf - (inf - x) * exp(-dt / tau), as written by the
    # reference (openclsim.cl:358-364)
    trhs = rl[s][1].rhs()
    if precision == DOUBLE_PRECISION:
        E = w.fn('exp') + '(-dt / ' + tau + ')'
    else:
        if isinstance(trhs, ex.Divide) and isinstance(
                trhs[0], ex.Number) and trhs[0].eval() == 1:
            # tau = 1 / X (the form the RL conversion writes), so
            # -dt / tau = -dt * X without two reciprocals
            rate = w.ex(trhs[1])
        else:
            rate = '__builtin_amdgcn_rcpf(' + tau + ')'
        # exp(-dt X) = 2^((-dt log2(e)) X): the first factor is the same
        # for every cell of a step, so one multiply and one v_exp_f32 per
        # gate; within 3 ulp of exp (OpenCL's bound) for dt X <= 1, and
        # below 1e-8 in absolute terms beyond
        E = ('__builtin_amdgcn_exp2f((-dt * 0x1.715476p+0f) * ' + rate
             + ')')
    return ('y[' + str(i) + '] = ' + inf + ' - (' + inf + ' - ' + x + ') * '
            + E + ';')


def diffusion_dependents(cm):
    """(variables, states): the intermediaries whose value depends on the
    diffusion current (with the variable bound to it), and the states whose
    derivative does. Everything else of a cell step can be evaluated before
    the neighbours' potentials of the step are known."""
    dep = set(v for v, what in cm.bound.items() if what == 'idiff')
    dstates = set()
    for var in cm.order:
        hit = False
        for r in var.rhs().references():
            v = r.var()
            if isinstance(r, ex.Derivative):
                hit = hit or v in dstates
            else:
                hit = hit or v in dep
        if hit:
            (dstates if var.is_state() else dep).add(var)
    return dep, dstates


def generate_split_cell_function(cm, precision, native_maths=False):
    """Returns the HIP source of ``mk_cell_a<LOG>`` and ``mk_cell_b<LOG>``,
    the cell step of ``mk_cell`` in two parts with the same operations:
    part A evaluates every equation, logged quantity and state update that
    does not depend on the diffusion current, and keeps what part B reads
    in ``carry`` (``MK_NCARRY`` values); part B, given the diffusion
    current, evaluates the rest. A persistent driver runs part A of all its
    cells before it waits for the neighbours' potentials of the step, and
    part B after. ``mk_qpart[k]`` says which part sets logged quantity k."""
    import re
    w = cell_writer(cm, precision, native_maths)
    tab = '    '
    dep, dstates = diffusion_dependents(cm)
    fam = []
    if precision == SINGLE_PRECISION:
        fam, members = exp_families(cm, w)
        w.set_exp_members(members)
    a_lines, b_lines = [], []
    declared = []
    for i, s in enumerate(cm.states):
        a_lines.append('const Real ' + _cname(s, 'x_') + ' = y[' + str(i)
                       + '];')
        declared.append(_cname(s, 'x_'))
    for var in cm.per_cell_constants:
        a_lines.append('const Real ' + _cname(var, 'c_') + ' = '
                       + w.ex(var.rhs()) + ';')
        declared.append(_cname(var, 'c_'))
    for line in fam:
        a_lines.append(line)
        declared.append(line.split()[2])
    for var in cm.order:
        if var.is_state():
            if var in cm.rl_states:
                continue
            name, into = _cname(var, 'd_'), (b_lines if var in dstates
                                               else a_lines)
        else:
            name, into = _cname(var), (b_lines if var in dep else a_lines)
        line = 'const Real ' + name + ' = ' + w.ex(var.rhs()) + ';'
        into.append(line)
        if into is a_lines:
            declared.append(name)
    qa, qb, qpart = [], [], []
    for k, var in enumerate(cm.quantities):
        line = 'q[' + str(k) + '] = ' + w.ex(ex.Name(var)) + ';'
        in_b = var in dep
        (qb if in_b else qa).append(line)
        qpart.append(1 if in_b else 0)
    ua, ub = [], []
    # a state that part B advances still holds its value at t in y when
    # part B starts (part A does not write it): B reads it there
    in_y = {}
    for i, s in enumerate(cm.states):
        if s in cm.rl_states:
            inf, tau = cm.rl_states[s]
            in_b = inf in dep or tau in dep
        else:
            in_b = s in dstates
        (ub if in_b else ua).append(_state_update(cm, w, precision, i, s))
        if in_b:
            in_y[_cname(s, 'x_')] = i
    # what part B reads of part A's values
    b_text = '\n'.join(b_lines + qb + ub)
    used = set(re.findall(r'[A-Za-z_][A-Za-z0-9_]*', b_text))
    from_y = [n for n in declared if n in used and n in in_y]
    carry = [n for n in declared if n in used and n not in in_y]
    out = ['#define MK_NCARRY ' + str(max(1, len(carry))),
           '__device__ constexpr int mk_qpart[' + str(max(1, len(qpart)))
           + '] = {' + ', '.join(str(x) for x in (qpart or [0])) + '};']
    sig = ('Real* __restrict__ y, const Real t, const Real dt, '
           'const Real pace,\n' + tab)
    out.append('template <bool LOG>')
    out.append('__device__ __forceinline__ void mk_cell_a(')
    out.append(tab + sig + 'const Real* __restrict__ fv, '
               'Real* __restrict__ q, Real* __restrict__ carry)')
    out.append('{')
    out.append(tab + '(void)t; (void)dt; (void)pace; (void)fv; (void)q;'
               ' (void)carry;')
    out += [tab + x for x in a_lines]
    if qa:
        out.append(tab + 'if (LOG) {')
        out += [2 * tab + x for x in qa]
        out.append(tab + '}')
    out += [tab + x for x in ua]
    out += [tab + 'carry[' + str(j) + '] = ' + n + ';'
            for j, n in enumerate(carry)]
    out.append('}')
    out.append('template <bool LOG>')
    out.append('__device__ __forceinline__ void mk_cell_b(')
    out.append(tab + sig + 'const Real idiff, const Real* __restrict__ fv, '
               'Real* __restrict__ q,\n' + tab
               + 'const Real* __restrict__ carry)')
    out.append('{')
    out.append(tab + '(void)t; (void)dt; (void)pace; (void)idiff; (void)fv;'
               ' (void)q; (void)carry;')
    out += [tab + 'const Real ' + n + ' = y[' + str(in_y[n]) + '];'
            for n in from_y]
    out += [tab + 'const Real ' + n + ' = carry[' + str(j) + '];'
            for j, n in enumerate(carry)]
    out += [tab + x for x in b_lines]
    if qb:
        out.append(tab + 'if (LOG) {')
        out += [2 * tab + x for x in qb]
        out.append(tab + '}')
    out += [tab + x for x in ub]
    out.append('}')
    return '\n'.join(out) + '\n'


_FLOAT_HELPERS = '''
/* exp(x) in single precision: p = x * log2(e) rounded, e = its rounding
 * error (exact through fma, plus the tail of log2(e)), then
 * exp(x) = 2^p * 2^e ~ 2^p * (1 + e * ln 2). One v_exp_f32 and five FMA-class
 * operations; within 2 ulp (OpenCL allows 3). Where 2^p is already 0 or
 * inf (underflow, overflow, x = -inf or +inf) it is the result: the
 * correction would turn inf into NaN when e < 0, and e is NaN for an
 * infinite x. */
__device__ __forceinline__ float mk_expf(float x)
{
    const float p = x * 0x1.715476p+0f;
    const float e = __builtin_fmaf(x, 0x1.4ae0c0p-26f,
                                   __builtin_fmaf(x, 0x1.715476p+0f, -p));
    const float r = __builtin_amdgcn_exp2f(p);
    const float y = __builtin_fmaf(r, e * 0x1.62e430p-1f, r);
    return (r == 0.0f || r == __builtin_inff()) ? r : y;
}

/* native_maths: 2^(x * log2(e)) directly, like OpenCL's native_exp */
__device__ __forceinline__ float mk_expf_native(float x)
{
    return __builtin_amdgcn_exp2f(x * 0x1.715476p+0f);
}
__device__ __forceinline__ float mk_powif(float x, int n)
{
    float r = 1.0f, b = x;
    const bool neg = n < 0;
    unsigned int u = (unsigned int)(neg ? -n : n);
    while (u) {
        if (u & 1u) r = r * b;
        u >>= 1;
        if (u) b = b * b;
    }
    return neg ? 1.0f / r : r;
}
'''


# Cells per thread of the persistent reverse sheet (MK_ACPT in the kernel
# source; the host library reads it from mkh_sim_config.adj_cpt)
ADJ_SHEET_CPT_CHOICES = (2, 4, 5, 6, 8)


def adjoint_sheet_cpt(precision):
    """Cells per thread of the persistent reverse sheet (tiles of
    64 x 4 cpt cells). Every tile is co-resident, so a CU holds its share of
    the sheet (4,096 cells of 1024^2) and mu of those cells in LDS (32 B
    per cell in single precision) whatever the tile size; the tile size
    only sets how many workgroups share it. 8: two workgroups of 73 KiB
    per CU, and the LR1991 reverse step takes 241 registers per lane
    without spills. 4 would give four workgroups per CU (four waves per
    SIMD) but caps the step at 128 registers, where it spills 68; 5 gives
    three per CU, whose 768 tiles per device are fewer than the 832 a
    1024^2 sheet needs."""
    return 8


def adjoint_sheet_groups(cm, precision, cpt=None):
    """Workgroups per CU (at most 4) that the LDS of the persistent reverse
    sheet leaves room for: mk_adj_sheet keeps mu of its 64 x 4 cpt tile and
    abar with a one-cell ring in LDS. 0 if one does not fit into a CU's
    160 KiB."""
    if cpt is None:
        cpt = adjoint_sheet_cpt(precision)
    w = 8 if precision == DOUBLE_PRECISION else 4
    th = 4 * cpt
    lds = len(cm.states) * th * 64 * w + (th + 2) * 66 * w + 64 + 4
    return min(4, (160 * 1024) // lds)


def adjoint_waves_default(cm, precision, adjoint_precision=None):
    """Waves per SIMD the reverse kernels are compiled for (1: no limit).
    The reverse step waits on its loads at the start of every cell; one
    more wave per SIMD hides that latency better than the few registers
    the limit spills cost, for cell models of up to 12 states with a
    single-precision VJP (LR1991: 135 registers unbounded, 128 and 6
    spilled at 4 waves)."""
    vjp = precision if adjoint_precision is None else adjoint_precision
    if vjp != DOUBLE_PRECISION and len(cm.states) <= 12:
        return 4
    return 1


def sheet_cpt(precision):
    """Cells per thread of the persistent sheet driver (4: a tile is 64 x 16
    cells for a 256-thread workgroup)."""
    return 4


def sheet_waves(precision):
    """Waves per SIMD the sheet kernel's register budget must allow: 4 in
    single precision (128 VGPRs, 4096 cells per CU, so 256 CUs hold 1024^2),
    2 in double (256 VGPRs; sheets up to half that size, larger ones use the
    grid driver)."""
    return 2 if precision == DOUBLE_PRECISION else 4


def grid_waves(precision):
    """Waves per SIMD the grid driver's register budget must allow (0: no
    limit, the compiler's choice). Double precision: 5, where the LR1991
    step takes 98 registers unbounded (4 waves) and 96 with two spilled;
    on the 1024^2 sheet 1.04e10 cell-steps/s unbounded, 1.11e10 at 5 and
    1.08e10 at 6 (24 spilled), bit-identical. With the branch-free exp the
    step takes 90 registers unbounded (5 waves), 72 at 6 or 7 waves, 64
    with spills at 8: 1.24e10 at 5, 6 and 7, 1.22e10 at 8; with its
    polynomial in fused multiply-adds, 92 unbounded, 74 at 6: 1.547e10 at
    5 and 6, 1.535e10 at 7 (spilled). Single precision needs no bound (38
    registers, 8 waves)."""
    return 5 if precision == DOUBLE_PRECISION else 0


def sheet_enabled(precision):
    return True


def chain_cpt(precision):
    """Cells per thread of the persistent chain driver: one-wave tiles of up
    to 64 * cpt cells. On 256 LR1991 cables of 50-2000 cells (256,157
    cells), single precision runs at 5.06e10 / 5.87e10 / 5.75e10
    cell-steps/s with 1 / 2 / 4 cells per thread and double precision at
    1.14e10 / 1.00e10 / 0.66e10 (the double-precision step holds twice the
    registers per cell, so one cell per thread leaves room for more
    waves)."""
    return 1 if precision == DOUBLE_PRECISION else 2


def chain_waves(precision):
    """Waves per SIMD the chain kernel's register budget must allow: 5 in
    single precision (so that 5 one-wave tiles of 64 cells per SIMD, 327,680
    cells, are co-resident on 256 CUs), 3 in double."""
    return 3 if precision == DOUBLE_PRECISION else 5


CHAIN_CPT_CHOICES = (1, 2, 4)


SHEET_CPT_CHOICES = (1, 2, 4, 8)


def generate_kernel(cm, precision, coupling, block_threads=256,
                    native_maths=False, adjoint=False, tile_cpt=None,
                    tile_waves=None, adjoint_precision=None,
                    adjoint_waves=None, chain_cells=None, chain_budget=None,
                    adjoint_cpt=None, grid_budget=None):
    """Returns the complete HIP source for hiprtc. ``tile_cpt`` (cells
    per thread of the persistent sheet: tile height 4 x cpt) and
    ``tile_waves`` (waves per SIMD its register budget allows) default to
    :func:`sheet_cpt` and :func:`sheet_waves`. With ``adjoint``, the
    adjoint kernels are appended; ``adjoint_precision`` is the precision of
    the cell VJP's arithmetic (default: ``precision``) and
    ``adjoint_waves`` the waves per SIMD the reverse kernels' register
    budget allows (0: no limit; default: :func:`adjoint_waves`).
    ``chain_cells`` and ``chain_budget`` set the chain driver's cells per
    thread and waves per SIMD (defaults :func:`chain_cpt`,
    :func:`chain_waves`), ``adjoint_cpt`` the persistent reverse sheet's
    cells per thread (default :func:`adjoint_sheet_cpt`), ``grid_budget``
    the waves per SIMD of the grid driver's register budget (default
    :func:`grid_waves`; 0: no limit)."""
    if block_threads % 64 or not 64 <= block_threads <= 1024:
        raise ValueError('block_threads must be a multiple of 64 in'
                         ' [64, 1024].')
    cpt = sheet_cpt(precision) if tile_cpt is None else int(tile_cpt)
    waves = sheet_waves(precision) if tile_waves is None \
        else int(tile_waves)
    if cpt not in SHEET_CPT_CHOICES:
        raise ValueError('sheet_cpt must be one of '
                         + str(SHEET_CPT_CHOICES) + '.')
    if not 1 <= waves <= 8:
        raise ValueError('sheet_waves must be in [1, 8].')
    ccpt = chain_cpt(precision) if chain_cells is None else int(chain_cells)
    cwaves = chain_waves(precision) if chain_budget is None \
        else int(chain_budget)
    if ccpt not in CHAIN_CPT_CHOICES:
        raise ValueError('chain_cpt must be one of '
                         + str(CHAIN_CPT_CHOICES) + '.')
    if not 1 <= cwaves <= 8:
        raise ValueError('chain_waves must be in [1, 8].')
    vm = cm.model.label('membrane_potential')
    i_vm = vm.index() if (vm is not None and vm.is_state()) else 0
    bcells = 2048 if precision == DOUBLE_PRECISION else 4096
    real = 'double' if precision == DOUBLE_PRECISION else 'float'
    acpt = adjoint_sheet_cpt(precision) if adjoint_cpt is None \
        else int(adjoint_cpt)
    if acpt not in ADJ_SHEET_CPT_CHOICES:
        raise ValueError('adjoint_cpt must be one of '
                         + str(ADJ_SHEET_CPT_CHOICES) + '.')
    asg = adjoint_sheet_groups(cm, precision, acpt)
    gw = grid_waves(precision) if grid_budget is None else int(grid_budget)
    if not 0 <= gw <= 8:
        raise ValueError('grid_waves must be in [0, 8].')
    # workgroups per CU for that many waves per SIMD (4 SIMDs per CU)
    gblocks = max(1, (gw * 256) // int(block_threads)) if gw else 0
    parts = [
        '// Generated by myokit_amd for model: '
        + str(cm.model.name()).replace('\n', ' '),
        '#define MK_NS ' + str(len(cm.states)),
        '#define MK_NF ' + str(len(cm.fields)),
        '#define MK_NQ ' + str(len(cm.quantities)),
        '#define MK_IVM ' + str(i_vm),
        '#define MK_COUPLING ' + str(int(coupling)),
        '#define MK_BT ' + str(int(block_threads)),
        '#define MK_BCELLS ' + str(bcells),
        '#define MK_SCPT ' + str(cpt),
        '#define MK_SWAVES ' + str(waves),
        '#define MK_CCPT ' + str(ccpt),
        '#define MK_CWAVES ' + str(cwaves),
        '#define MK_ACPT ' + str(acpt),
        '#define MK_GBLOCKS ' + str(gblocks),
        '#define MK_ADJ_SHEET ' + ('1' if adjoint and int(coupling) == 1
                                   and asg > 0
                                   and adjoint_precision in (None, precision)
                                   else '0'),
        '#define MK_ADJ_SWAVES ' + str(max(1, asg)),
        '#define MK_ADJ_WAVES ' + str(max(1, int(
            adjoint_waves_default(cm, precision, adjoint_precision)
            if adjoint_waves is None else adjoint_waves))),
        'typedef ' + real + ' Real;',
        _read('mk_math.h'),
        _read('mk_args.h'),
        _FLOAT_HELPERS,
        generate_cell_function(cm, precision, native_maths),
        generate_split_cell_function(cm, precision, native_maths),
        _read('mk_drivers.hip'),
    ]
    if adjoint:
        from ._adjointgen import generate_vjp_function
        parts.append(generate_vjp_function(cm, precision, adjoint_precision))
        parts.append(_read('mk_adjoint.hip'))
    return '\n'.join(parts)


def compile_options(precision):
    """hiprtc options. Contraction is disabled in double precision so that
    results match the CPU oracle bit for bit. Single precision fuses a
    multiply and an add only within one source expression
    (``-ffp-contract=on``): with ``fast`` the compiler also fuses across
    statements, differently in each kernel, and the drivers (and runs split
    into several launches) stop giving identical results. Packed FP32
    instructions are disabled."""
    opts = ['-O3', '-std=c++17',
            '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']
    if precision == DOUBLE_PRECISION:
        opts.append('-ffp-contract=off')
    else:
        opts.append('-ffp-contract=on')
    return ' '.join(opts)
