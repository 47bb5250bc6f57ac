"""
Expression tree for model equations.

This is the small subset of myokit's expression system that the tissue engine
needs: the node types of ``myokit/_expressions.py`` (``Number:812``,
``Name:999``, ``Derivative:1159``, prefix/infix operators ``:1474-2095``,
functions ``:2095-2935``, conditions ``:2935-3267``), evaluation with IEEE-754
semantics (no exceptions: ``1/0`` gives ``inf`` as it does on the device), and
symbolic partial derivatives following ``Expression.diff``
(``_expressions.py:243-336``; the rules themselves are ordinary calculus).

Operator semantics follow the reference's C writers
(``formats/ansic/_ewriter.py:49-121``): ``a // b`` is ``floor(a / b)``,
``a % b`` is ``a - b * floor(a / b)``, ``log(a, b)`` is ``log(a) / log(b)``.
"""
import math

# Binding powers, used to decide where brackets are needed in code()
BP_OR = 10
BP_AND = 20
BP_NOT = 25
BP_COMPARE = 30
BP_SUM = 40
BP_PRODUCT = 50
BP_PREFIX = 60
BP_POWER = 70
BP_ATOM = 100


# ---------------------------------------------------------------------------
# IEEE-754 helpers (Python's math raises where hardware returns inf/nan)
# ---------------------------------------------------------------------------
INF = float('inf')
NAN = float('nan')


def ieee_div(a, b):
    try:
        return a / b
    except ZeroDivisionError:
        if a != a or a == 0:
            return NAN
        return math.copysign(INF, a) * math.copysign(1.0, b)


def ieee_exp(a):
    try:
        return math.exp(a)
    except OverflowError:
        return INF


def ieee_log(a):
    if a != a or a < 0:
        return NAN
    if a == 0:
        return -INF
    if a == INF:
        return INF
    return math.log(a)


def ieee_log10(a):
    if a != a or a < 0:
        return NAN
    if a == 0:
        return -INF
    if a == INF:
        return INF
    return math.log10(a)


def ieee_sqrt(a):
    if a != a or a < 0:
        return NAN
    return math.sqrt(a)


def ieee_pow(a, b):
    try:
        r = math.pow(a, b)
    except OverflowError:
        r = INF
    except (ValueError, ZeroDivisionError):
        if a == 0 and b < 0:
            return INF
        return NAN
    return r


def _guard(f, a):
    try:
        return f(a)
    except (ValueError, OverflowError):
        return NAN


# ---------------------------------------------------------------------------
# Base class
# ---------------------------------------------------------------------------
class Expression:
    """Base class of every node in an expression tree."""
    _bp = BP_ATOM
    _rep = ''

    def __init__(self, operands=()):
        self._ops = tuple(operands)
        for op in self._ops:
            if not isinstance(op, Expression):
                raise TypeError(
                    'Operands must be Expression objects, got '
                    + type(op).__name__)

    # Sequence interface over operands
    def __iter__(self):
        return iter(self._ops)

    def __getitem__(self, i):
        return self._ops[i]

    def __len__(self):
        return len(self._ops)

    # Structural identity
    def _key(self):
        return (type(self).__name__,) + tuple(op._key() for op in self._ops)

    def __eq__(self, other):
        if not isinstance(other, Expression):
            return False
        return self._key() == other._key()

    def __ne__(self, other):
        return not self.__eq__(other)

    def __hash__(self):
        return hash(self._key())

    def __str__(self):
        return self.code()

    def __repr__(self):
        return '<' + type(self).__name__ + ' ' + self.code() + '>'

    # Tree walking
    def walk(self):
        """Yields every node in this tree, depth-first, parents first."""
        stack = [self]
        while stack:
            e = stack.pop()
            yield e
            stack.extend(reversed(e._ops))

    def references(self):
        """Returns the set of LhsExpressions used directly in this tree."""
        return set(e for e in self.walk() if isinstance(e, LhsExpression))

    def depends_on(self, lhs, deep=False):
        """True if ``lhs`` appears in this tree (optionally through the
        right-hand sides of referenced variables)."""
        seen = set()
        stack = [self]
        while stack:
            e = stack.pop()
            for x in e.walk():
                if x == lhs:
                    return True
                if deep and isinstance(x, Name) and not isinstance(
                        x._value, str):
                    var = x._value
                    if var not in seen and not var.is_state() and \
                            var.rhs() is not None:
                        seen.add(var)
                        stack.append(var.rhs())
        return False

    def is_literal(self):
        """True if this tree contains no variable references."""
        return not any(isinstance(e, LhsExpression) for e in self.walk())

    def is_conditional(self):
        return any(isinstance(e, (If, Piecewise)) for e in self.walk())

    # Evaluation
    def eval(self, subst=None):
        """Evaluates this expression with IEEE-754 double semantics.

        ``subst`` maps LhsExpressions (or Variables) to values; names that are
        not substituted are evaluated through their variable's RHS.
        """
        s = {}
        if subst:
            for k, v in subst.items():
                if not isinstance(k, Expression):
                    k = Name(k)
                s[k] = float(v)
        return self._eval(s)

    def _eval(self, subst):
        raise NotImplementedError

    # Cloning with substitution
    def clone(self, subst=None, expand=False, retain=None):
        """Returns a deep copy, with substitutions ``{lhs: expr}`` applied.

        With ``expand=True`` every name is replaced by its variable's RHS,
        except for states, bound variables and names in ``retain``.
        """
        if subst and self in subst:
            return subst[self]
        if expand and isinstance(self, Name) and not isinstance(
                self._value, str):
            var = self._value
            keep = retain and (self in retain or var in retain)
            if not (keep or var.is_state() or var.is_bound()
                    or var.rhs() is None):
                return var.rhs().clone(subst, expand, retain)
        return self._rebuild(
            [op.clone(subst, expand, retain) for op in self._ops])

    def _rebuild(self, ops):
        return type(self)(*ops)

    # Derivatives
    def diff(self, lhs, independent_states=True):
        """Returns the partial derivative of this expression w.r.t. ``lhs``.

        Names of intermediary variables are differentiated through their
        right-hand side (chain rule). With ``independent_states=True`` states
        other than ``lhs`` count as independent (derivative zero). Bound
        variables always have derivative zero.
        """
        if not isinstance(lhs, (Name, Derivative)):
            raise ValueError(
                'Partial derivatives can only be taken with respect to a'
                ' Name or Derivative.')
        d = self._diff(lhs, _DiffContext(bool(independent_states)))
        return Number(0) if d is None else d

    def local_diff(self, ref):
        """Partial derivative w.r.t. the LhsExpression ``ref``, treating every
        other reference as an independent symbol (no chain rule)."""
        d = self._diff(ref, None)
        return Number(0) if d is None else d

    def _diff(self, lhs, ctx):
        raise NotImplementedError

    # Code
    def code(self):
        raise NotImplementedError

    def _bracket(self, op, right=False):
        s = op.code()
        if op._bp < self._bp or (right and op._bp == self._bp):
            return '(' + s + ')'
        return s


class _DiffContext:
    """State for chain-rule differentiation through variable RHSs."""

    def __init__(self, independent_states):
        self.independent_states = independent_states
        self.cache = {}


# Helpers that treat None as a symbolic zero
def _add(a, b):
    if a is None:
        return b
    if b is None:
        return a
    return Plus(a, b)


def _sub(a, b):
    if b is None:
        return a
    if a is None:
        return PrefixMinus(b)
    return Minus(a, b)


def _mul(a, b):
    if a is None or b is None:
        return None
    return Multiply(a, b)


def _neg(a):
    return None if a is None else PrefixMinus(a)


# ---------------------------------------------------------------------------
# Atoms
# ---------------------------------------------------------------------------
class Number(Expression):
    """A literal number (units are parsed but not kept)."""
    _bp = BP_ATOM

    def __init__(self, value, unit=None):
        super().__init__()
        if isinstance(value, Number):
            value = value._value
        self._value = float(value)
        self._unit = unit

    def _key(self):
        v = self._value
        return ('Number', v if v == v else 'nan')

    def _rebuild(self, ops):
        return Number(self._value, self._unit)

    def code(self):
        # repr() gives the shortest string that round-trips to the double
        return repr(self._value)

    def eval(self, subst=None):
        return self._value

    def _eval(self, subst):
        return self._value

    def _diff(self, lhs, ctx):
        return None

    def value(self):
        return self._value


class LhsExpression(Expression):
    """Base class for Name and Derivative."""

    def var(self):
        return self._value


class Name(LhsExpression):
    """A reference to a variable (or, before resolution, a name string)."""

    def __init__(self, value):
        super().__init__()
        self._value = value

    def _key(self):
        v = self._value
        if isinstance(v, str):
            return ('Name', v)
        return ('Name', id(v))

    def _rebuild(self, ops):
        return Name(self._value)

    def code(self):
        v = self._value
        if isinstance(v, str):
            return v
        return v.qname()

    def _eval(self, subst):
        try:
            return subst[self]
        except KeyError:
            pass
        v = self._value
        if isinstance(v, str):
            raise ValueError('Unresolved name <' + v + '>.')
        rhs = v.rhs()
        if rhs is None:
            raise ValueError('Variable <' + v.qname() + '> has no RHS.')
        if v.is_state():
            # A state's value is its initial value
            return float(v.initial_value_float())
        return rhs._eval(subst)

    def _diff(self, lhs, ctx):
        if self == lhs:
            return Number(1)
        if ctx is None or isinstance(self._value, str):
            return None
        var = self._value
        if var.is_bound() or var.is_state():
            return None
        # Chain rule through the RHS of an intermediary or constant
        key = (id(var), lhs._key())
        if key not in ctx.cache:
            ctx.cache[key] = var.rhs()._diff(lhs, ctx)
        return ctx.cache[key]


class Derivative(LhsExpression):
    """The time derivative ``dot(x)`` of a state variable."""

    def __init__(self, op):
        if not isinstance(op, Name):
            raise TypeError('Derivative operand must be a Name.')
        super().__init__((op,))
        self._value = op._value

    def _key(self):
        return ('Derivative',) + self._ops[0]._key()

    def var(self):
        return self._ops[0]._value

    def code(self):
        return 'dot(' + self._ops[0].code() + ')'

    def _eval(self, subst):
        try:
            return subst[self]
        except KeyError:
            pass
        return self.var().rhs()._eval(subst)

    def _diff(self, lhs, ctx):
        if self == lhs:
            return Number(1)
        if ctx is None:
            return None
        return self.var().rhs()._diff(lhs, ctx)


# ---------------------------------------------------------------------------
# Prefix operators
# ---------------------------------------------------------------------------
class PrefixPlus(Expression):
    _bp = BP_PREFIX
    _rep = '+'

    def __init__(self, op):
        super().__init__((op,))

    def code(self):
        return '+' + self._bracket(self._ops[0])

    def _eval(self, subst):
        return +self._ops[0]._eval(subst)

    def _diff(self, lhs, ctx):
        return self._ops[0]._diff(lhs, ctx)


class PrefixMinus(Expression):
    _bp = BP_PREFIX
    _rep = '-'

    def __init__(self, op):
        super().__init__((op,))

    def code(self):
        s = self._bracket(self._ops[0])
        if s.startswith('-'):
            s = '(' + s + ')'
        return '-' + s

    def _eval(self, subst):
        return -self._ops[0]._eval(subst)

    def _diff(self, lhs, ctx):
        return _neg(self._ops[0]._diff(lhs, ctx))


# ---------------------------------------------------------------------------
# Infix operators
# ---------------------------------------------------------------------------
class InfixExpression(Expression):
    def __init__(self, left, right):
        super().__init__((left, right))

    def code(self):
        return (self._bracket(self._ops[0]) + ' ' + self._rep + ' '
                + self._bracket(self._ops[1], right=True))


class Plus(InfixExpression):
    _bp = BP_SUM
    _rep = '+'

    def _eval(self, subst):
        return self._ops[0]._eval(subst) + self._ops[1]._eval(subst)

    def _diff(self, lhs, ctx):
        return _add(self._ops[0]._diff(lhs, ctx),
                    self._ops[1]._diff(lhs, ctx))


class Minus(InfixExpression):
    _bp = BP_SUM
    _rep = '-'

    def _eval(self, subst):
        return self._ops[0]._eval(subst) - self._ops[1]._eval(subst)

    def _diff(self, lhs, ctx):
        return _sub(self._ops[0]._diff(lhs, ctx),
                    self._ops[1]._diff(lhs, ctx))


class Multiply(InfixExpression):
    _bp = BP_PRODUCT
    _rep = '*'

    def _eval(self, subst):
        return self._ops[0]._eval(subst) * self._ops[1]._eval(subst)

    def _diff(self, lhs, ctx):
        a, b = self._ops
        return _add(_mul(a._diff(lhs, ctx), b), _mul(a, b._diff(lhs, ctx)))


class Divide(InfixExpression):
    _bp = BP_PRODUCT
    _rep = '/'

    def _eval(self, subst):
        return ieee_div(self._ops[0]._eval(subst), self._ops[1]._eval(subst))

    def _diff(self, lhs, ctx):
        # d(a/b) = da / b - (a / b) * db / b
        a, b = self._ops
        da, db = a._diff(lhs, ctx), b._diff(lhs, ctx)
        t1 = None if da is None else Divide(da, b)
        t2 = None if db is None else Divide(Multiply(self, db), b)
        return _sub(t1, t2)


class Quotient(InfixExpression):
    """Integer division rounding towards minus infinity."""
    _bp = BP_PRODUCT
    _rep = '//'

    def _eval(self, subst):
        q = ieee_div(self._ops[0]._eval(subst), self._ops[1]._eval(subst))
        return float(math.floor(q)) if math.isfinite(q) else q

    def _diff(self, lhs, ctx):
        return None


class Remainder(InfixExpression):
    """``a % b = a - b * floor(a / b)``."""
    _bp = BP_PRODUCT
    _rep = '%'

    def _eval(self, subst):
        a = self._ops[0]._eval(subst)
        b = self._ops[1]._eval(subst)
        q = ieee_div(a, b)
        if not math.isfinite(q):
            return NAN
        return a - b * math.floor(q)

    def _diff(self, lhs, ctx):
        a, b = self._ops
        da, db = a._diff(lhs, ctx), b._diff(lhs, ctx)
        return _sub(da, _mul(db, Floor(Divide(a, b))))


class Power(InfixExpression):
    _bp = BP_POWER
    _rep = '^'

    def code(self):
        return (self._bracket(self._ops[0], right=True) + '^'
                + self._bracket(self._ops[1], right=True))

    def _eval(self, subst):
        return ieee_pow(self._ops[0]._eval(subst), self._ops[1]._eval(subst))

    def _diff(self, lhs, ctx):
        a, b = self._ops
        da, db = a._diff(lhs, ctx), b._diff(lhs, ctx)
        if db is None:
            if da is None:
                return None
            # b * a^(b - 1) * da
            return Multiply(Multiply(b, Power(a, Minus(b, Number(1)))), da)
        # a^b * (db * log(a) + b * da / a)
        inner = _add(Multiply(db, Log(a)),
                     None if da is None else Divide(Multiply(b, da), a))
        return Multiply(self, inner)


# ---------------------------------------------------------------------------
# Functions
# ---------------------------------------------------------------------------
class Function(Expression):
    _fname = None
    _nargs = (1,)

    def __init__(self, *ops):
        if len(ops) not in self._nargs:
            raise ValueError(
                f'Function {self._fname}() takes {self._nargs} arguments,'
                f' got {len(ops)}.')
        super().__init__(ops)

    def code(self):
        return self._fname + '(' + ', '.join(o.code() for o in self._ops) \
            + ')'


class Sqrt(Function):
    _fname = 'sqrt'

    def _eval(self, subst):
        return ieee_sqrt(self._ops[0]._eval(subst))

    def _diff(self, lhs, ctx):
        da = self._ops[0]._diff(lhs, ctx)
        if da is None:
            return None
        return Divide(da, Multiply(Number(2), self))


class Exp(Function):
    _fname = 'exp'

    def _eval(self, subst):
        return ieee_exp(self._ops[0]._eval(subst))

    def _diff(self, lhs, ctx):
        return _mul(self, self._ops[0]._diff(lhs, ctx))


class Log(Function):
    _fname = 'log'
    _nargs = (1, 2)

    def _eval(self, subst):
        a = ieee_log(self._ops[0]._eval(subst))
        if len(self._ops) == 1:
            return a
        return ieee_div(a, ieee_log(self._ops[1]._eval(subst)))

    def _diff(self, lhs, ctx):
        if len(self._ops) == 1:
            da = self._ops[0]._diff(lhs, ctx)
            return None if da is None else Divide(da, self._ops[0])
        # log(a, b) = log(a) / log(b)
        return Divide(Log(self._ops[0]), Log(self._ops[1]))._diff(lhs, ctx)


class Log10(Function):
    _fname = 'log10'

    def _eval(self, subst):
        return ieee_log10(self._ops[0]._eval(subst))

    def _diff(self, lhs, ctx):
        da = self._ops[0]._diff(lhs, ctx)
        if da is None:
            return None
        return Divide(da, Multiply(self._ops[0], Log(Number(10))))


class Sin(Function):
    _fname = 'sin'

    def _eval(self, subst):
        return _guard(math.sin, self._ops[0]._eval(subst))

    def _diff(self, lhs, ctx):
        return _mul(Cos(self._ops[0]), self._ops[0]._diff(lhs, ctx))


class Cos(Function):
    _fname = 'cos'

    def _eval(self, subst):
        return _guard(math.cos, self._ops[0]._eval(subst))

    def _diff(self, lhs, ctx):
        return _neg(_mul(Sin(self._ops[0]), self._ops[0]._diff(lhs, ctx)))


class Tan(Function):
    _fname = 'tan'

    def _eval(self, subst):
        return _guard(math.tan, self._ops[0]._eval(subst))

    def _diff(self, lhs, ctx):
        da = self._ops[0]._diff(lhs, ctx)
        if da is None:
            return None
        return Divide(da, Power(Cos(self._ops[0]), Number(2)))


class ASin(Function):
    _fname = 'asin'

    def _eval(self, subst):
        return _guard(math.asin, self._ops[0]._eval(subst))

    def _diff(self, lhs, ctx):
        a = self._ops[0]
        da = a._diff(lhs, ctx)
        if da is None:
            return None
        return Divide(da, Sqrt(Minus(Number(1), Power(a, Number(2)))))


class ACos(Function):
    _fname = 'acos'

    def _eval(self, subst):
        return _guard(math.acos, self._ops[0]._eval(subst))

    def _diff(self, lhs, ctx):
        a = self._ops[0]
        da = a._diff(lhs, ctx)
        if da is None:
            return None
        return PrefixMinus(
            Divide(da, Sqrt(Minus(Number(1), Power(a, Number(2))))))


class ATan(Function):
    _fname = 'atan'

    def _eval(self, subst):
        return _guard(math.atan, self._ops[0]._eval(subst))

    def _diff(self, lhs, ctx):
        a = self._ops[0]
        da = a._diff(lhs, ctx)
        if da is None:
            return None
        return Divide(da, Plus(Number(1), Power(a, Number(2))))


class Floor(Function):
    _fname = 'floor'

    def _eval(self, subst):
        a = self._ops[0]._eval(subst)
        return float(math.floor(a)) if math.isfinite(a) else a

    def _diff(self, lhs, ctx):
        return None


class Ceil(Function):
    _fname = 'ceil'

    def _eval(self, subst):
        a = self._ops[0]._eval(subst)
        return float(math.ceil(a)) if math.isfinite(a) else a

    def _diff(self, lhs, ctx):
        return None


class Abs(Function):
    _fname = 'abs'

    def _eval(self, subst):
        return abs(self._ops[0]._eval(subst))

    def _diff(self, lhs, ctx):
        a = self._ops[0]
        da = a._diff(lhs, ctx)
        if da is None:
            return None
        return If(MoreEqual(a, Number(0)), da, PrefixMinus(da))


class If(Function):
    """``if(condition, then, else)``."""
    _fname = 'if'
    _nargs = (3,)

    def condition(self):
        return self._ops[0]

    def _eval(self, subst):
        if self._ops[0]._eval(subst):
            return self._ops[1]._eval(subst)
        return self._ops[2]._eval(subst)

    def _diff(self, lhs, ctx):
        dt = self._ops[1]._diff(lhs, ctx)
        de = self._ops[2]._diff(lhs, ctx)
        if dt is None and de is None:
            return None
        return If(self._ops[0], dt or Number(0), de or Number(0))


class Piecewise(Function):
    """``piecewise(c1, v1, c2, v2, ..., else)``."""
    _fname = 'piecewise'

    def __init__(self, *ops):
        if len(ops) < 3 or len(ops) % 2 == 0:
            raise ValueError(
                'Piecewise requires an odd number of arguments (at least 3).')
        Expression.__init__(self, ops)

    def conditions(self):
        return self._ops[0:-1:2]

    def pieces(self):
        return self._ops[1::2]

    def _eval(self, subst):
        ops = self._ops
        for i in range(0, len(ops) - 1, 2):
            if ops[i]._eval(subst):
                return ops[i + 1]._eval(subst)
        return ops[-1]._eval(subst)

    def _diff(self, lhs, ctx):
        ds = [p._diff(lhs, ctx) for p in self.pieces()]
        if all(d is None for d in ds):
            return None
        ops = []
        for c, d in zip(self.conditions(), ds[:-1]):
            ops.extend([c, d or Number(0)])
        ops.append(ds[-1] or Number(0))
        return Piecewise(*ops)


# ---------------------------------------------------------------------------
# Conditions
# ---------------------------------------------------------------------------
class Condition:
    """Mixin for expressions that evaluate to a truth value."""


class Not(Expression, Condition):
    _bp = BP_NOT

    def __init__(self, op):
        super().__init__((op,))

    def code(self):
        return 'not ' + self._bracket(self._ops[0])

    def _eval(self, subst):
        return not self._ops[0]._eval(subst)

    def _diff(self, lhs, ctx):
        return None


class _Compare(InfixExpression, Condition):
    _bp = BP_COMPARE

    def _diff(self, lhs, ctx):
        return None


class Equal(_Compare):
    _rep = '=='

    def _eval(self, subst):
        return self._ops[0]._eval(subst) == self._ops[1]._eval(subst)


class NotEqual(_Compare):
    _rep = '!='

    def _eval(self, subst):
        return self._ops[0]._eval(subst) != self._ops[1]._eval(subst)


class More(_Compare):
    _rep = '>'

    def _eval(self, subst):
        return self._ops[0]._eval(subst) > self._ops[1]._eval(subst)


class Less(_Compare):
    _rep = '<'

    def _eval(self, subst):
        return self._ops[0]._eval(subst) < self._ops[1]._eval(subst)


class MoreEqual(_Compare):
    _rep = '>='

    def _eval(self, subst):
        return self._ops[0]._eval(subst) >= self._ops[1]._eval(subst)


class LessEqual(_Compare):
    _rep = '<='

    def _eval(self, subst):
        return self._ops[0]._eval(subst) <= self._ops[1]._eval(subst)


class And(InfixExpression, Condition):
    _bp = BP_AND
    _rep = 'and'

    def _eval(self, subst):
        return bool(self._ops[0]._eval(subst)) and bool(
            self._ops[1]._eval(subst))

    def _diff(self, lhs, ctx):
        return None


class Or(InfixExpression, Condition):
    _bp = BP_OR
    _rep = 'or'

    def _eval(self, subst):
        return bool(self._ops[0]._eval(subst)) or bool(
            self._ops[1]._eval(subst))

    def _diff(self, lhs, ctx):
        return None


# Function name -> class, used by the parser
FUNCTIONS = {
    'sqrt': Sqrt, 'exp': Exp, 'log': Log, 'log10': Log10,
    'sin': Sin, 'cos': Cos, 'tan': Tan,
    'asin': ASin, 'acos': ACos, 'atan': ATan,
    'floor': Floor, 'ceil': Ceil, 'abs': Abs,
    'if': If, 'piecewise': Piecewise,
}
