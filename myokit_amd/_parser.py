"""
Parser for the subset of myokit's ``mmt`` format used by cell models.

Supported (the syntax of ``myokit/_bin/example.mmt`` and of the models in
``myokit/tests/data``, described in ``myokit/_parsing.py``): the ``[[model]]``
header (meta data and initial values), ``[component]`` sections with meta
data, ``use a.b as c`` aliases, definitions ``x = ...`` and ``dot(x) = ...``
with trailing ``in [unit]``, ``bind``, ``label`` and ``: description``,
nested variables by indentation, ``\\`` and bracket line continuation,
``\"\"\"`` multi-line meta text, numbers with units, and the ``[[protocol]]``
table. ``[[script]]`` sections are skipped. Units are parsed and discarded:
they have no effect on simulation.

The implementation is a line classifier plus a Pratt expression parser; it
does not follow the structure of the reference tokenizer.
"""
import os
import re

from . import _expressions as ex
from ._err import ParseError
from ._model import Model

# Operator precedence (left binding power) for the Pratt parser
_INFIX = {
    'or': (10, ex.Or), 'and': (20, ex.And),
    '==': (30, ex.Equal), '!=': (30, ex.NotEqual), '<': (30, ex.Less),
    '>': (30, ex.More), '<=': (30, ex.LessEqual), '>=': (30, ex.MoreEqual),
    '+': (40, ex.Plus), '-': (40, ex.Minus),
    '*': (50, ex.Multiply), '/': (50, ex.Divide), '//': (50, ex.Quotient),
    '%': (50, ex.Remainder),
    '^': (70, ex.Power),
}
_PREFIX_BP = 60
_NOT_BP = 25

_TOKEN = re.compile(r'''
    (?P<ws>\s+)
  | (?P<num>(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)
  | (?P<unit>\[[^\]]*\])
  | (?P<name>[a-zA-Z_]\w*(?:\.[a-zA-Z_]\w*)*)
  | (?P<op>//|==|!=|<=|>=|[-+*/^%<>(),])
''', re.VERBOSE)

_DEF = re.compile(r'^(dot\(\s*([a-zA-Z_]\w*)\s*\)|[a-zA-Z_]\w*)\s*=(?!=)(.*)$')
_META = re.compile(r'^([a-zA-Z_][\w:\-]*)\s*:(.*)$')
_INIT = re.compile(
    r'^([a-zA-Z_]\w*(?:\.[a-zA-Z_]\w*)+)\s*=(?!=)\s*(.*)$')
_USE = re.compile(r'^use\s+(.*)$')


def _tokenize(text, line):
    toks = []
    pos = 0
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if m is None:
            raise ParseError('Unexpected character', line, pos,
                             repr(text[pos:pos + 10]))
        pos = m.end()
        kind = m.lastgroup
        if kind == 'ws':
            continue
        toks.append((kind, m.group(kind), m.start()))
    toks.append(('eof', None, pos))
    return toks


class _ExprParser:
    """Pratt parser over one tokenized line; stops at keywords."""

    STOP = ('in', 'bind', 'label')

    def __init__(self, toks, line):
        self.toks = toks
        self.i = 0
        self.line = line

    def peek(self):
        return self.toks[self.i]

    def next(self):
        t = self.toks[self.i]
        self.i += 1
        return t

    def expect(self, value):
        t = self.next()
        if t[1] != value:
            raise ParseError('Syntax error', self.line, t[2],
                             'Expected "' + value + '", got ' + repr(t[1]))
        return t

    def at_stop(self):
        t = self.peek()
        return t[0] == 'eof' or (t[0] == 'name' and t[1] in self.STOP)

    def parse(self, rbp=0):
        left = self.nud()
        while True:
            t = self.peek()
            if t[0] == 'eof' or t[1] in (')', ','):
                break
            if t[0] == 'name' and t[1] not in ('and', 'or'):
                break
            op = t[1]
            if op not in _INFIX:
                raise ParseError('Syntax error', self.line, t[2],
                                 'Unexpected ' + repr(op))
            bp, cls = _INFIX[op]
            if bp <= rbp:
                break
            self.next()
            # '^' is right-associative, everything else left-associative
            right = self.parse(bp - 1 if op == '^' else bp)
            left = cls(left, right)
        return left

    def nud(self):
        t = self.next()
        kind, val, pos = t
        if kind == 'num':
            n = ex.Number(float(val))
            if self.peek()[0] == 'unit':
                self.next()
            return n
        if kind == 'op' and val == '(':
            e = self.parse()
            self.expect(')')
            return e
        if kind == 'op' and val == '-':
            return ex.PrefixMinus(self.parse(_PREFIX_BP))
        if kind == 'op' and val == '+':
            return ex.PrefixPlus(self.parse(_PREFIX_BP))
        if kind == 'name':
            if val == 'not':
                return ex.Not(self.parse(_NOT_BP))
            if self.peek()[1] == '(':
                return self.call(val, pos)
            if val in ('and', 'or') or val in self.STOP:
                raise ParseError('Syntax error', self.line, pos,
                                 'Unexpected keyword ' + val)
            return ex.Name(val)
        raise ParseError('Syntax error', self.line, pos,
                         'Unexpected ' + repr(val))

    def call(self, fname, pos):
        self.expect('(')
        args = []
        if self.peek()[1] != ')':
            while True:
                args.append(self.parse())
                if self.peek()[1] == ',':
                    self.next()
                    continue
                break
        self.expect(')')
        if fname == 'dot':
            if len(args) != 1 or not isinstance(args[0], ex.Name):
                raise ParseError('Syntax error', self.line, pos,
                                 'dot() takes a single variable name')
            return ex.Derivative(args[0])
        cls = ex.FUNCTIONS.get(fname)
        if cls is None:
            raise ParseError('Unknown function', self.line, pos, fname)
        try:
            return cls(*args)
        except ValueError as e:
            raise ParseError('Invalid function call', self.line, pos, str(e))


def parse_expression(text, line=0):
    """Parses a single expression (names stay unresolved strings)."""
    p = _ExprParser(_tokenize(text, line), line)
    e = p.parse()
    if p.peek()[0] != 'eof':
        raise ParseError('Syntax error', line, p.peek()[2],
                         'Unexpected text after expression')
    return e


def _logical_lines(source):
    """Yields (line_number, indent, text, is_meta_text) after removing
    comments, joining continuations and multi-line strings."""
    lines = source.splitlines()
    i = 0
    n = len(lines)
    while i < n:
        raw = lines[i]
        lineno = i + 1
        i += 1
        stripped = raw.strip()
        if not stripped:
            continue
        indent = len(raw) - len(raw.lstrip(' \t'))
        # Multi-line meta strings
        m = _META.match(stripped)
        if m and not _DEF.match(stripped) and not _INIT.match(stripped):
            text = m.group(2).strip()
            if text.startswith('"""'):
                body = text[3:]
                while '"""' not in body and i < n:
                    body += '\n' + lines[i]
                    i += 1
                body = body.split('"""')[0]
                text = body.strip()
            yield lineno, indent, (m.group(1), text), True
            continue
        # Remove comment
        if '#' in stripped:
            stripped = stripped[:stripped.index('#')].rstrip()
            if not stripped:
                continue
        # Continuations: trailing backslash or open brackets
        while (stripped.endswith('\\') or
               stripped.count('(') > stripped.count(')')) and i < n:
            if stripped.endswith('\\'):
                stripped = stripped[:-1]
            nxt = lines[i]
            i += 1
            if '#' in nxt:
                nxt = nxt[:nxt.index('#')]
            stripped = stripped + ' ' + nxt.strip()
        yield lineno, indent, stripped, False


def parse(source):
    """Parses mmt source text; returns ``(model, protocol, script)``."""
    from ._protocol import Protocol
    model = protocol = None
    script = None
    section = None
    comp = None
    stack = []          # (indent, variable) for nesting
    inits = []          # (line, qname, expression)
    proto_rows = []
    lines = list(_logical_lines(source))
    for lineno, indent, text, is_meta in lines:
        if not is_meta and text.startswith('[[') and text.endswith(']]'):
            section = text[2:-2].strip()
            if section == 'model':
                model = Model()
            elif section == 'protocol':
                protocol = Protocol()
            elif section == 'script':
                script = ''
                break
            else:
                raise ParseError('Unknown segment', lineno, 0, text)
            comp = None
            continue
        if not is_meta and text.startswith('[') and text.endswith(']') \
                and section == 'model':
            name = text[1:-1].strip()
            try:
                comp = model.add_component(name)
            except Exception as e:
                raise ParseError('Invalid component', lineno, 0, str(e))
            stack = []
            continue

        if section == 'protocol':
            if is_meta:
                raise ParseError('Unexpected meta data in protocol', lineno)
            proto_rows.append((lineno, text))
            continue
        if section != 'model':
            raise ParseError('Text outside of a segment', lineno, 0, text)

        if comp is None:
            # Model header: meta data and initial values
            if is_meta:
                model.meta[text[0]] = text[1]
                continue
            m = _INIT.match(text)
            if not m:
                raise ParseError('Invalid initial value', lineno, 0, text)
            inits.append((lineno, m.group(1), parse_expression(
                m.group(2), lineno)))
            continue

        # Pop nesting levels at or deeper than this indent
        while stack and stack[-1][0] >= indent:
            stack.pop()
        owner = stack[-1][1] if stack else None

        if is_meta:
            key, value = text
            (owner if owner is not None else comp).meta[key] = value
            continue

        m = _USE.match(text)
        if m and owner is None:
            for part in m.group(1).split(','):
                part = part.strip()
                mm = re.match(r'^([\w.]+)(?:\s+as\s+(\w+))?$', part)
                if not mm:
                    raise ParseError('Invalid use statement', lineno, 0, text)
                target = mm.group(1)
                alias = mm.group(2) or target.split('.')[-1]
                comp._pending_alias = getattr(comp, '_pending_alias', [])
                comp._pending_alias.append((lineno, alias, target))
            continue

        m = _DEF.match(text)
        if m:
            lhs, state_name, rest = m.group(1), m.group(2), m.group(3)
            name = state_name or lhs
            desc = None
            if ':' in rest:
                rest, desc = rest.split(':', 1)
                desc = desc.strip()
            parent = owner if owner is not None else comp
            try:
                var = parent.add_variable(name)
            except Exception as e:
                raise ParseError('Invalid variable', lineno, 0, str(e))
            toks = _tokenize(rest, lineno)
            p = _ExprParser(toks, lineno)
            var._rhs = p.parse()
            var._is_state_pending = bool(state_name)
            _parse_properties(p, var, lineno)
            if desc:
                var.meta['desc'] = desc
            stack.append((indent, var))
            continue

        # Property line for the owner
        if owner is None:
            raise ParseError('Unexpected line', lineno, 0, text)
        p = _ExprParser(_tokenize(text, lineno), lineno)
        _parse_properties(p, owner, lineno)

    if model is not None:
        _finish_model(model, inits)
    if protocol is not None:
        _finish_protocol(protocol, proto_rows)
    return model, protocol, script


def _parse_properties(p, var, lineno):
    """Parses trailing ``in [unit]``, ``bind x`` and ``label x``."""
    while True:
        t = p.next()
        if t[0] == 'eof':
            return
        if t[0] == 'name' and t[1] == 'in':
            u = p.next()
            if u[0] != 'unit':
                raise ParseError('Expected unit', lineno, u[2])
            var.set_unit(u[1][1:-1])
        elif t[0] == 'name' and t[1] == 'bind':
            b = p.next()
            if b[0] != 'name':
                raise ParseError('Expected binding label', lineno, b[2])
            var._pending_binding = b[1]
        elif t[0] == 'name' and t[1] == 'label':
            b = p.next()
            if b[0] != 'name':
                raise ParseError('Expected label', lineno, b[2])
            var._pending_label = b[1]
        else:
            raise ParseError('Syntax error', lineno, t[2],
                             'Unexpected ' + repr(t[1]))


def _finish_model(model, inits):
    # States, in the order of the initial values
    pending_states = {}
    for c in model.components():
        for v in c.variables(deep=True):
            if getattr(v, '_is_state_pending', False):
                pending_states[v.qname()] = v
    for lineno, qname, expr in inits:
        v = pending_states.pop(qname, None)
        if v is None:
            raise ParseError('Initial value for unknown or non-state variable',
                             lineno, 0, qname)
        v._is_state = True
        v._initial_value = expr
        model._states.append(v)
    if pending_states:
        raise ParseError('Missing initial value', 0, 0,
                         ', '.join(sorted(pending_states)))
    # Aliases, bindings, labels
    for c in model.components():
        for lineno, alias, target in getattr(c, '_pending_alias', []):
            try:
                c.add_alias(alias, model.get(target))
            except KeyError:
                raise ParseError('Unknown variable in use statement', lineno,
                                 0, target)
        if hasattr(c, '_pending_alias'):
            del c._pending_alias
        for v in c.variables(deep=True):
            if hasattr(v, '_is_state_pending'):
                del v._is_state_pending
            b = getattr(v, '_pending_binding', None)
            if b is not None:
                del v._pending_binding
                v.set_binding(b)
            lab = getattr(v, '_pending_label', None)
            if lab is not None:
                del v._pending_label
                v.set_label(lab)
    model.validate()


def _finish_protocol(protocol, rows):
    last_end = 0.0
    for lineno, text in rows:
        parts = text.split()
        if len(parts) != 5:
            raise ParseError('Protocol rows need 5 values', lineno, 0, text)
        try:
            level = float(parts[0])
            start = last_end if parts[1] == 'next' else float(parts[1])
            duration = float(parts[2])
            period = float(parts[3])
            mult = int(float(parts[4]))
        except ValueError:
            raise ParseError('Invalid protocol row', lineno, 0, text)
        protocol.schedule(level, start, duration, period, mult)
        last_end = start + duration


def parse_model(source):
    """Parses an mmt model from a string."""
    model, _, _ = parse(source)
    if model is None:
        raise ParseError('No [[model]] segment found', 0)
    return model


def parse_protocol(source):
    _, protocol, _ = parse(source)
    return protocol


_EXAMPLE = os.path.join(os.path.dirname(__file__), 'data', 'example.mmt')


def load(path):
    """Loads ``(model, protocol, script)`` from an mmt file; ``'example'``
    loads the bundled Luo-Rudy 1991 model."""
    if path == 'example':
        path = _EXAMPLE
    with open(path, 'r') as f:
        return parse(f.read())


def load_model(path):
    model, _, _ = load(path)
    if model is None:
        raise ParseError('No [[model]] segment found', 0)
    return model


def load_protocol(path):
    return load(path)[1]
