"""
Model, Component and Variable: the parts of myokit's model API
(``myokit/_model_api.py``) that the tissue engine uses.

Behaviour followed (not code): states are ordered by the initial-value lines of
the ``[[model]]`` header; names resolve in the scope of the variable (its own
children, then its ancestors' children, then the component, then ``use``
aliases, then fully qualified names); bound variables are inputs; a constant
is a non-state, unbound variable whose RHS only references constants
(``Variable.is_constant``, ``_model_api.py:4425-4456``); ``solvable_order``
(``:3100-3317``) is reproduced as one global topological order, which yields
the same values because every equation is a pure function of its inputs.
"""
import math
import re
from collections import OrderedDict

from . import _expressions as ex
from ._err import (
    CyclicalDependencyError, IllegalReferenceError, IntegrityError,
    NumericalError, MissingRhsError)

_NAME = re.compile(r'^[a-zA-Z_]\w*$')


class _VarOwner:
    """Shared code of Component and Variable: an ordered set of children."""

    def __init__(self):
        self._vars = OrderedDict()

    def add_variable(self, name):
        if not _NAME.match(name):
            raise ValueError('Invalid variable name <' + str(name) + '>.')
        if name in self._vars:
            raise IntegrityError(
                'Duplicate variable name <' + name + '> in <' + self.qname()
                + '>.')
        v = Variable(self, name)
        self._vars[name] = v
        self.model()._invalidate()
        return v

    def add_variable_allow_renaming(self, name):
        base, i = name, 1
        while self._has_in_scope(name):
            name = base + '_' + str(i)
            i += 1
        return self.add_variable(name)

    def _has_in_scope(self, name):
        # A new child must not shadow or duplicate a name visible here
        if name in self._vars:
            return True
        p = self.parent()
        return p._has_in_scope(name) if isinstance(p, _VarOwner) else False

    def variables(self, deep=False):
        for v in self._vars.values():
            yield v
            if deep:
                yield from v.variables(deep=True)

    def count_variables(self):
        return len(self._vars)

    def has_variable(self, name):
        return name in self._vars

    def get(self, name):
        parts = name.split('.')
        x = self
        for p in parts:
            x = x._vars[p]
        return x

    def move_variable(self, variable, new_parent, new_name=None):
        """Moves a child variable (with its nested children) to another
        component or variable (``Component.move_variable``,
        ``myokit/_model_api.py``). References to it keep pointing at it: the
        model's names are resolved to variables first."""
        if variable._parent is not self:
            raise ValueError('Variable <' + variable.qname()
                             + '> is not a child of <' + self.qname() + '>.')
        self.model().validate()
        name = new_name or variable.name()
        if not _NAME.match(name):
            raise ValueError('Invalid variable name <' + str(name) + '>.')
        if name in new_parent._vars:
            raise IntegrityError(
                'Duplicate variable name <' + name + '> in <'
                + new_parent.qname() + '>.')
        del self._vars[variable.name()]
        variable._parent = new_parent
        variable._name = name
        new_parent._vars[name] = variable
        self.model()._invalidate()


class Variable(_VarOwner):
    """A model variable (state, constant, intermediary or bound input)."""

    def __init__(self, parent, name):
        super().__init__()
        self._parent = parent
        self._name = name
        self._rhs = None
        self._is_state = False
        self._initial_value = None
        self._binding = None
        self._label = None
        self._unit = None
        self.meta = OrderedDict()
        self._index = None
        self._uname = None

    # Naming and hierarchy
    def name(self):
        return self._name

    def qname(self, hide=None):
        return self._parent.qname() + '.' + self._name

    def uname(self):
        """A unique C identifier for this variable (after
        Model.create_unique_names())."""
        return self._uname or self.qname().replace('.', '_')

    def parent(self, kind=None):
        if kind is None:
            return self._parent
        p = self._parent
        while p is not None and not isinstance(p, kind):
            p = p.parent() if not isinstance(p, Model) else None
        return p

    def component(self):
        p = self._parent
        while isinstance(p, Variable):
            p = p._parent
        return p

    def model(self):
        return self.component().model()

    def is_nested(self):
        return isinstance(self._parent, Variable)

    # RHS
    def rhs(self):
        return self._rhs

    def set_rhs(self, rhs):
        """Sets the right-hand side: an Expression, a number, or the text
        of an expression (names resolved when the model is validated)."""
        if isinstance(rhs, str):
            try:
                rhs = ex.Number(float(rhs))
            except ValueError:
                from ._parser import parse_expression
                rhs = parse_expression(rhs)
        elif rhs is not None and not isinstance(rhs, ex.Expression):
            rhs = ex.Number(rhs)
        self._rhs = rhs
        self.model()._invalidate()

    def lhs(self):
        n = ex.Name(self)
        return ex.Derivative(n) if self._is_state else n

    # States
    def is_state(self):
        return self._is_state

    def promote(self, initial_value=0):
        if self._is_state:
            raise IntegrityError('Variable is already a state.')
        if self.is_nested():
            raise IntegrityError('Nested variables cannot be states.')
        self._is_state = True
        if not isinstance(initial_value, ex.Expression):
            initial_value = ex.Number(initial_value)
        self._initial_value = initial_value
        self.model()._add_state(self)

    def demote(self):
        self._is_state = False
        self._initial_value = None
        self.model()._remove_state(self)

    def index(self):
        if not self._is_state:
            raise ValueError('Only states have an index.')
        return self.model()._states.index(self)

    def initial_value(self, as_float=False):
        if as_float:
            return self.initial_value_float()
        return self._initial_value

    def initial_value_float(self):
        return float(self._initial_value.eval())

    def set_initial_value(self, value):
        if not isinstance(value, ex.Expression):
            value = ex.Number(value)
        self._initial_value = value

    # Bindings and labels
    def binding(self):
        return self._binding

    def set_binding(self, binding):
        m = self.model()
        if self._binding is not None:
            m._bindings.pop(self._binding, None)
        if binding is not None:
            if binding in m._bindings and m._bindings[binding] is not self:
                raise IntegrityError(
                    'Duplicate binding <' + binding + '>.')
            if self._is_state:
                raise IntegrityError('States cannot be bound.')
            m._bindings[binding] = self
        self._binding = binding
        m._invalidate()

    def is_bound(self):
        return self._binding is not None

    def label(self):
        return self._label

    def set_label(self, label):
        m = self.model()
        if self._label is not None:
            m._labels.pop(self._label, None)
        if label is not None:
            if label in m._labels and m._labels[label] is not self:
                raise IntegrityError('Duplicate label <' + label + '>.')
            m._labels[label] = self
        self._label = label

    def unit(self):
        return self._unit

    def set_unit(self, unit):
        self._unit = unit

    # Classification (valid after Model.validate())
    def is_constant(self):
        return self.model()._classify()[self] in ('literal', 'calculated')

    def is_literal(self):
        return self.model()._classify()[self] == 'literal'

    def is_intermediary(self):
        return self.model()._classify()[self] == 'inter'

    def __repr__(self):
        return '<Variable(' + self.qname() + ')>'

    def __str__(self):
        return self.qname()


class Component(_VarOwner):
    """A component: a named group of variables plus ``use`` aliases."""

    def __init__(self, model, name):
        super().__init__()
        self._model = model
        self._name = name
        self._aliases = OrderedDict()
        self.meta = OrderedDict()

    def name(self):
        return self._name

    def qname(self):
        return self._name

    def parent(self, kind=None):
        return self._model

    def model(self):
        return self._model

    def _has_in_scope(self, name):
        return name in self._vars or name in self._aliases

    def add_alias(self, name, variable):
        if name in self._vars or name in self._aliases:
            raise IntegrityError(
                'Alias <' + name + '> clashes with an existing name in <'
                + self._name + '>.')
        self._aliases[name] = variable

    def alias_for(self, variable):
        for k, v in self._aliases.items():
            if v is variable:
                return k
        return None

    def __repr__(self):
        return '<Component(' + self._name + ')>'


class Model:
    """A cell model: components, variables, states, bindings and labels."""

    def __init__(self, name=None):
        self._components = OrderedDict()
        self._states = []
        self._bindings = OrderedDict()
        self._labels = OrderedDict()
        self.meta = OrderedDict()
        if name is not None:
            self.meta['name'] = name
        self._cache = None

    # Structure
    def name(self):
        return self.meta.get('name')

    def model(self):
        return self

    def add_component(self, name):
        if not _NAME.match(name):
            raise ValueError('Invalid component name <' + str(name) + '>.')
        if name in self._components:
            raise IntegrityError('Duplicate component <' + name + '>.')
        c = Component(self, name)
        self._components[name] = c
        return c

    def add_component_allow_renaming(self, name):
        base, i = name, 1
        while name in self._components:
            name = base + '_' + str(i)
            i += 1
        return self.add_component(name)

    def components(self):
        return iter(self._components.values())

    def has_component(self, name):
        return name in self._components

    def __getitem__(self, name):
        """The component with the given name (``Model.__getitem__``,
        ``myokit/_model_api.py``)."""
        return self._components[name]

    def __contains__(self, name):
        return name in self._components

    def get(self, name, class_=None):
        """Returns the component or variable with the given qname."""
        if isinstance(name, (Variable, Component)):
            name = name.qname()
        parts = name.split('.')
        try:
            x = self._components[parts[0]]
            for p in parts[1:]:
                x = x._vars[p]
        except KeyError:
            raise KeyError('Component or variable not found: "' + name + '"')
        if class_ is not None and not isinstance(x, class_):
            raise KeyError('Object <' + name + '> is not a '
                           + class_.__name__)
        return x

    def variables(self, state=None, const=None, inter=None, bound=None,
                  deep=False):
        """Iterates over variables, optionally filtered by class."""
        cls = self._classify()
        for c in self._components.values():
            for v in c.variables(deep=deep):
                k = cls[v]
                if state is not None and (k == 'state') != state:
                    continue
                if const is not None and (
                        k in ('literal', 'calculated')) != const:
                    continue
                if inter is not None and (k == 'inter') != inter:
                    continue
                if bound is not None and (k == 'bound') != bound:
                    continue
                yield v

    def count_variables(self, deep=False):
        return sum(1 for _ in self.variables(deep=deep))

    # States
    def _add_state(self, var):
        self._states.append(var)
        self._invalidate()

    def _remove_state(self, var):
        self._states.remove(var)
        self._invalidate()

    def states(self):
        return iter(list(self._states))

    def count_states(self):
        return len(self._states)

    def reorder_state(self, order):
        if set(order) != set(self._states) or len(order) != len(self._states):
            raise ValueError('Order must contain every state exactly once.')
        self._states = list(order)

    def initial_values(self, as_floats=False):
        if as_floats:
            return [v.initial_value_float() for v in self._states]
        return [v._initial_value for v in self._states]

    def state(self):
        return self.initial_values(as_floats=True)

    def set_state(self, values):
        values = list(values)
        if len(values) != len(self._states):
            raise ValueError('Wrong number of state values.')
        for v, x in zip(self._states, values):
            v.set_initial_value(float(x))

    # Bindings and labels
    def binding(self, label):
        return self._bindings.get(label)

    def bindings(self):
        return list(self._bindings.items())

    def label(self, label):
        return self._labels.get(label)

    def labels(self):
        return list(self._labels.items())

    def time(self):
        return self._bindings.get('time')

    def set_value(self, qname, value):
        self.get(qname).set_rhs(ex.Number(float(value)))

    # Analysis
    def _invalidate(self):
        self._cache = None

    def _analysis(self):
        if self._cache is None:
            self._cache = _Analysis(self)
        return self._cache

    def _classify(self):
        return self._analysis().kind

    def validate(self):
        """Checks the model and raises an IntegrityError if it is invalid."""
        self._invalidate()
        self._analysis()

    def is_valid(self):
        try:
            self.validate()
            return True
        except IntegrityError:
            return False

    def solvable_order(self):
        """Returns every non-constant equation's variable in an order in
        which it can be evaluated (states and bound variables are inputs).
        Derivatives of states are listed as their state variable."""
        return list(self._analysis().order)

    def constants_order(self):
        """Returns every constant in evaluation order."""
        return list(self._analysis().const_order)

    def has_interdependent_components(self):
        return len(self.component_cycles()) > 0

    def component_cycles(self):
        """Returns cycles in the component dependency graph (components
        whose equations depend on each other through non-state variables)."""
        a = self._analysis()
        edges = {c: set() for c in self._components.values()}
        for v, deps in a.deps.items():
            cv = v.component()
            for d in deps:
                cd = d.component()
                if cd is not cv:
                    edges[cv].add(cd)
        # Tarjan's strongly connected components
        index, low, onstack, stack, out = {}, {}, set(), [], []
        counter = [0]

        def strong(v):
            index[v] = low[v] = counter[0]
            counter[0] += 1
            stack.append(v)
            onstack.add(v)
            for w in edges[v]:
                if w not in index:
                    strong(w)
                    low[v] = min(low[v], low[w])
                elif w in onstack:
                    low[v] = min(low[v], index[w])
            if low[v] == index[v]:
                scc = []
                while True:
                    w = stack.pop()
                    onstack.discard(w)
                    scc.append(w)
                    if w is v:
                        break
                if len(scc) > 1:
                    out.append(scc)

        for c in edges:
            if c not in index:
                strong(c)
        return out

    def create_unique_names(self):
        """Gives every variable a unique C identifier (``uname``)."""
        used = set()
        for c in self._components.values():
            for v in c.variables(deep=True):
                base = re.sub(r'\W', '_', v.qname())
                name, i = base, 1
                while name in used:
                    name = base + '_' + str(i)
                    i += 1
                used.add(name)
                v._uname = name

    def evaluate_derivatives(self, state=None, inputs=None,
                             precision=None, ignore_errors=True):
        """Evaluates the state derivatives at ``state`` with bound inputs
        ``inputs`` (a dict ``{binding_label: value}``), in double precision.
        """
        state = self.state() if state is None else list(state)
        if len(state) != len(self._states):
            raise ValueError('State has the wrong size.')
        subst = {}
        for v, x in zip(self._states, state):
            subst[ex.Name(v)] = float(x)
        inputs = inputs or {}
        for label, var in self._bindings.items():
            if label in inputs:
                subst[ex.Name(var)] = float(inputs[label])
        values = {}
        for var in self._analysis().const_order + self._analysis().order:
            n = ex.Name(var)
            if n in subst and not var.is_state():
                continue
            try:
                x = var.rhs()._eval(subst)
            except Exception as e:      # pragma: no cover
                if not ignore_errors:
                    raise NumericalError(str(e))
                x = float('nan')
            if not ignore_errors and not math.isfinite(x):
                raise NumericalError(
                    'Non-finite value for <' + var.qname() + '>.')
            if var.is_state():
                values[var] = x
            else:
                subst[n] = x
        return [values[v] for v in self._states]

    def clone(self):
        """Returns a deep copy of this model."""
        m = Model()
        m.meta = OrderedDict(self.meta)
        vmap = {}

        def copy_vars(src, dst):
            for v in src._vars.values():
                w = Variable(dst, v._name)
                dst._vars[v._name] = w
                w._unit = v._unit
                w.meta = OrderedDict(v.meta)
                vmap[v] = w
                copy_vars(v, w)

        for c in self._components.values():
            d = m.add_component(c._name)
            d.meta = OrderedDict(c.meta)
            copy_vars(c, d)
        for c in self._components.values():
            d = m._components[c._name]
            for a, v in c._aliases.items():
                d._aliases[a] = vmap[v]

        def remap(e):
            if e is None:
                return None
            subst = {}
            for r in e.references():
                if isinstance(r, ex.Derivative):
                    subst[r] = ex.Derivative(ex.Name(vmap[r.var()]))
                elif not isinstance(r._value, str):
                    subst[r] = ex.Name(vmap[r._value])
            return e.clone(subst=subst)

        for v, w in vmap.items():
            w._rhs = remap(v._rhs)
            w._is_state = v._is_state
            w._initial_value = remap(v._initial_value)
            w._binding = v._binding
            w._label = v._label
            w._uname = v._uname
        m._states = [vmap[v] for v in self._states]
        m._bindings = OrderedDict(
            (k, vmap[v]) for k, v in self._bindings.items())
        m._labels = OrderedDict((k, vmap[v]) for k, v in self._labels.items())
        return m

    def code(self):
        """Returns an mmt-like text representation (for debugging)."""
        out = ['[[model]]']
        for k, v in self.meta.items():
            out.append(k + ': ' + v.replace('\n', ' '))
        for s in self._states:
            out.append(s.qname() + ' = ' + s._initial_value.code())
        for c in self._components.values():
            out.append('')
            out.append('[' + c._name + ']')
            for a, v in c._aliases.items():
                out.append('use ' + v.qname() + ' as ' + a)

            def write(v, indent):
                lhs = 'dot(' + v._name + ')' if v._is_state else v._name
                out.append('    ' * indent + lhs + ' = ' + (
                    v._rhs.code() if v._rhs is not None else '?'))
                if v._binding:
                    out.append('    ' * (indent + 1) + 'bind ' + v._binding)
                if v._label:
                    out.append('    ' * (indent + 1) + 'label ' + v._label)
                for w in v._vars.values():
                    write(w, indent + 1)

            for v in c._vars.values():
                write(v, 0)
        return '\n'.join(out) + '\n'

    def format_state(self, state=None, precision=None):
        state = self.state() if state is None else state
        w = max(len(s.qname()) for s in self._states) if self._states else 0
        return '\n'.join(s.qname().ljust(w) + ' = ' + repr(float(x))
                         for s, x in zip(self._states, state))

    def format_state_derivatives(self, state=None, derivatives=None,
                                 precision=None):
        state = self.state() if state is None else state
        if derivatives is None:
            derivatives = self.evaluate_derivatives(state)
        w = max(len(s.qname()) for s in self._states) if self._states else 0
        return '\n'.join(
            s.qname().ljust(w) + ' = ' + repr(float(x)).ljust(24)
            + '   dot = ' + repr(float(d))
            for s, x, d in zip(self._states, state, derivatives))


class _Analysis:
    """Resolves names, checks the model and orders its equations."""

    def __init__(self, model):
        self.model = model
        allvars = []
        for c in model._components.values():
            allvars.extend(c.variables(deep=True))

        # Resolve string names
        for v in allvars:
            if v._rhs is None:
                raise MissingRhsError(
                    'No RHS set for <' + v.qname() + '>.')
            v._rhs = _resolve(v, v._rhs)
        for s in model._states:
            if s._initial_value is None:
                raise IntegrityError(
                    'No initial value set for state <' + s.qname() + '>.')
            s._initial_value = _resolve(s, s._initial_value)

        # Kind of each variable and dependencies
        kind = {}
        for v in allvars:
            if v._is_state:
                kind[v] = 'state'
            elif v._binding is not None:
                kind[v] = 'bound'
        deps = {}
        for v in allvars:
            d = set()
            for r in v._rhs.references():
                if isinstance(r, ex.Derivative):
                    d.add(r.var())
                else:
                    w = r._value
                    if not w._is_state:
                        d.add(w)
            if v._binding is not None:
                d = set()
            deps[v] = d
        self.deps = deps

        # Topological sort (Kahn), states' derivatives are nodes too
        order = []
        indeg = {v: 0 for v in allvars}
        users = {v: [] for v in allvars}
        for v, ds in deps.items():
            for d in ds:
                indeg[v] += 1
                users[d].append(v)
        ready = [v for v in allvars if indeg[v] == 0]
        ready.reverse()
        while ready:
            v = ready.pop()
            order.append(v)
            for u in users[v]:
                indeg[u] -= 1
                if indeg[u] == 0:
                    ready.append(u)
        if len(order) != len(allvars):
            bad = [v.qname() for v in allvars if indeg[v] > 0]
            raise CyclicalDependencyError(
                'Cyclical dependency found involving: ' + ', '.join(bad))

        # Constants: unbound non-states whose references are all constants
        for v in order:
            if v in kind:
                continue
            literal = True
            const = True
            for r in v._rhs.references():
                if isinstance(r, ex.Derivative):
                    const = False
                    break
                w = r._value
                literal = False
                if kind.get(w) not in ('literal', 'calculated'):
                    const = False
                    break
            if const:
                kind[v] = 'literal' if literal else 'calculated'
            else:
                kind[v] = 'inter'
        self.kind = kind
        self.const_order = [v for v in order
                            if kind[v] in ('literal', 'calculated')]
        self.order = [v for v in order if kind[v] in ('inter', 'state')]


def _resolve(var, expr):
    """Replaces string names in ``expr`` by variables, using the scope of
    ``var``."""
    subst = {}
    for r in expr.references():
        if isinstance(r, ex.Derivative):
            n = r[0]
            if isinstance(n._value, str):
                w = _lookup(var, n._value)
                if not w._is_state:
                    raise IntegrityError(
                        'dot() used on non-state <' + w.qname() + '>.')
                subst[r] = ex.Derivative(ex.Name(w))
        elif isinstance(r._value, str):
            subst[r] = ex.Name(_lookup(var, r._value))
    if not subst:
        return expr
    return expr.clone(subst=subst)


def _lookup(var, name):
    # Own children, then ancestors' children, then component and aliases
    if '.' not in name:
        x = var
        while isinstance(x, Variable):
            if name in x._vars:
                return x._vars[name]
            x = x._parent
        comp = x
        if name in comp._vars:
            return comp._vars[name]
        if name in comp._aliases:
            return comp._aliases[name]
        raise IllegalReferenceError(
            'Unknown variable <' + name + '> referenced in <'
            + var.qname() + '>.')
    try:
        return var.model().get(name)
    except KeyError:
        raise IllegalReferenceError(
            'Unknown variable <' + name + '> referenced in <'
            + var.qname() + '>.')
