"""
CPU ORACLE -- test infrastructure, not part of the product.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module, and only as the checker / CPU baseline.

Writes a model's equations as plain C (``oracle_cell``: evaluate every
equation of one cell in dependency order, then the forward-Euler or
Rush-Larsen update, as in ``cell_step``, ``myokit/_sim/openclsim.cl:298-367``;
C semantics of ``myokit/formats/ansic/_ewriter.py:49-121``: floor-based
quotient and remainder, ternary conditionals, ``log(a)/log(b)``), compiles it
with ``oracle/mk_oracle.c`` using gcc (``-O2 -ffp-contract=off -fopenmp``)
into ``oracle/_ref/``-style build folders under ``oracle/_build``, and runs
it through ctypes.

This writer is separate from the product's HIP writer. It uses the same
portable math header (``myokit_amd/csrc/mk_math.h``) so that double-precision
results can be compared bit for bit, and the same model parser and RL
detection (those are pinned separately against the reference: golden
derivatives, golden inf/tau detection).
"""
import ctypes
import hashlib
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, '_build')


def cpu_flags():
    """gcc flags for the building CPU: ``-mfma`` where it has the FMA3
    instruction, so that the fused multiply-adds of ``mk_math.h`` compile
    to one instruction instead of a call to the C library's fma (the two
    round alike: both are correctly rounded)."""
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('flags'):
                    return ['-mfma'] if ' fma ' in line + ' ' else []
    except OSError:
        pass
    return []
MATH_H = os.path.join(ROOT, 'myokit_amd', 'csrc', 'mk_math.h')
DRIVER_C = os.path.join(HERE, 'mk_oracle.c')


# --------------------------------------------------------------------------
# Constant evaluation (host double precision, C library functions)
# --------------------------------------------------------------------------
def _c_div(a, b):
    try:
        return a / b
    except ZeroDivisionError:
        if a != a or a == 0:
            return float('nan')
        return math.copysign(float('inf'), a) * math.copysign(1.0, b)


def _c_call(f, *a):
    try:
        return f(*a)
    except OverflowError:
        return float('inf')
    except ValueError:
        return float('nan')


def const_eval(e, values):
    """Evaluates a constant expression; ``values`` maps variables to
    floats."""
    t = type(e).__name__
    ops = list(e)
    if t == 'Number':
        return e.eval()
    if t == 'Name':
        return values[e.var()]
    v = [const_eval(o, values) for o in ops] if t not in (
        'If', 'Piecewise') else None
    if t == 'PrefixPlus':
        return v[0]
    if t == 'PrefixMinus':
        return -v[0]
    if t == 'Plus':
        return v[0] + v[1]
    if t == 'Minus':
        return v[0] - v[1]
    if t == 'Multiply':
        return v[0] * v[1]
    if t == 'Divide':
        return _c_div(v[0], v[1])
    if t == 'Quotient':
        q = _c_div(v[0], v[1])
        return float(math.floor(q)) if math.isfinite(q) else q
    if t == 'Remainder':
        q = _c_div(v[0], v[1])
        return v[0] - v[1] * math.floor(q) if math.isfinite(q) else \
            float('nan')
    if t == 'Power':
        return _c_call(math.pow, v[0], v[1])
    if t == 'Sqrt':
        return _c_call(math.sqrt, v[0])
    if t == 'Exp':
        return _c_call(math.exp, v[0])
    if t == 'Log':
        if len(v) == 1:
            return _c_call(math.log, v[0])
        return _c_div(_c_call(math.log, v[0]), _c_call(math.log, v[1]))
    fn = {'Log10': math.log10, 'Sin': math.sin, 'Cos': math.cos,
          'Tan': math.tan, 'ASin': math.asin, 'ACos': math.acos,
          'ATan': math.atan, 'Abs': abs}
    if t in fn:
        return _c_call(fn[t], v[0])
    if t == 'Floor':
        return float(math.floor(v[0]))
    if t == 'Ceil':
        return float(math.ceil(v[0]))
    if t == 'If':
        c = const_eval(ops[0], values)
        return const_eval(ops[1] if c else ops[2], values)
    if t == 'Piecewise':
        for i in range(0, len(ops) - 1, 2):
            if const_eval(ops[i], values):
                return const_eval(ops[i + 1], values)
        return const_eval(ops[-1], values)
    cmp = {'Equal': lambda a, b: a == b, 'NotEqual': lambda a, b: a != b,
           'More': lambda a, b: a > b, 'Less': lambda a, b: a < b,
           'MoreEqual': lambda a, b: a >= b, 'LessEqual': lambda a, b: a <= b,
           'And': lambda a, b: bool(a) and bool(b),
           'Or': lambda a, b: bool(a) or bool(b)}
    if t in cmp:
        return cmp[t](v[0], v[1])
    if t == 'Not':
        return not v[0]
    raise ValueError('Cannot evaluate ' + t)


# --------------------------------------------------------------------------
# C writer
# --------------------------------------------------------------------------
class CWriter:
    FUN = {'Sqrt': 'sqrt', 'Exp': 'mk_exp', 'Log10': 'mk_log10',
           'Sin': 'mk_sin', 'Cos': 'mk_cos', 'Tan': 'mk_tan',
           'ASin': 'mk_asin', 'ACos': 'mk_acos', 'ATan': 'mk_atan',
           'Floor': 'floor', 'Ceil': 'ceil', 'Abs': 'fabs'}
    BIN = {'Plus': '+', 'Minus': '-', 'Multiply': '*', 'Divide': '/',
           'Equal': '==', 'NotEqual': '!=', 'More': '>', 'Less': '<',
           'MoreEqual': '>=', 'LessEqual': '<=', 'And': '&&', 'Or': '||'}

    def __init__(self, names, constants):
        self.names = names
        self.constants = constants

    @staticmethod
    def num(x):
        if math.isnan(x):
            return '(0.0/0.0)'
        if math.isinf(x):
            return '(1.0/0.0)' if x > 0 else '(-1.0/0.0)'
        return '(' + repr(float(x)) + ')'

    def w(self, e):
        t = type(e).__name__
        ops = list(e)
        if t == 'Number':
            return self.num(e.eval())
        if t in ('Name', 'Derivative'):
            if e in self.names:
                return self.names[e]
            return self.num(self.constants[e.var()])
        if t == 'PrefixPlus':
            return '(+' + self.w(ops[0]) + ')'
        if t == 'PrefixMinus':
            return '(-' + self.w(ops[0]) + ')'
        if t in self.BIN:
            return ('(' + self.w(ops[0]) + ' ' + self.BIN[t] + ' '
                    + self.w(ops[1]) + ')')
        if t == 'Quotient':
            return 'floor(' + self.w(ops[0]) + ' / ' + self.w(ops[1]) + ')'
        if t == 'Remainder':
            a, b = self.w(ops[0]), self.w(ops[1])
            return '(' + a + ' - ' + b + ' * floor(' + a + ' / ' + b + '))'
        if t == 'Power':
            return 'mk_pow(' + self.w(ops[0]) + ', ' + self.w(ops[1]) + ')'
        if t == 'Log':
            if len(ops) == 1:
                return 'mk_log(' + self.w(ops[0]) + ')'
            return ('(mk_log(' + self.w(ops[0]) + ') / mk_log('
                    + self.w(ops[1]) + '))')
        if t in self.FUN:
            return self.FUN[t] + '(' + self.w(ops[0]) + ')'
        if t == 'Not':
            return '(!' + self.w(ops[0]) + ')'
        if t == 'If':
            return ('(' + self.w(ops[0]) + ' ? ' + self.w(ops[1]) + ' : '
                    + self.w(ops[2]) + ')')
        if t == 'Piecewise':
            s = ''
            for i in range(0, len(ops) - 1, 2):
                s += '(' + self.w(ops[i]) + ' ? ' + self.w(ops[i + 1]) + ' : '
            return s + self.w(ops[-1]) + ')' * ((len(ops) - 1) // 2)
        raise ValueError('Cannot write ' + t)


def model_c_parts(model, bound, fields):
    """What both C writers of a model need: (writer, states, per-cell
    constants, solvable order, names). Constants that depend on no field
    are folded to numbers; those that do are evaluated per cell (``pc``)."""
    from myokit_amd import _expressions as ex
    states = list(model.states())
    fields = list(fields)
    fset = set(fields)
    names = {}
    for var, what in bound.items():
        names[ex.Name(var)] = {'time': 't', 'pace': 'pace',
                               'idiff': 'idiff'}[what]
    for i, s in enumerate(states):
        names[ex.Name(s)] = 'x' + str(i)
        names[ex.Derivative(ex.Name(s))] = 'dx' + str(i)
    for k, f in enumerate(fields):
        names[ex.Name(f)] = 'fv[' + str(k) + ']'
    constants = {}
    percell = []
    for var in model.constants_order():
        if var in fset:
            continue
        refs = [r.var() for r in var.rhs().references()]
        if any(r in fset or r in percell for r in refs):
            percell.append(var)
            names[ex.Name(var)] = 'pc' + str(len(percell) - 1)
            continue
        constants[var] = const_eval(var.rhs(), constants)
    order = model.solvable_order()
    k = 0
    for var in order:
        if not var.is_state():
            names[ex.Name(var)] = 'v' + str(k)
            k += 1
    return CWriter(names, constants), states, percell, order, names


def model_source(model, bound, rl_states, fields, quantities,
                 log_after_update=False):
    """C source of ``oracle_cell`` for a prepared model.

    Logged quantities are evaluated from the state at the step's time t,
    before the update (``openclsim.cl:298-367`` writes ``inter_log`` from
    the state it is about to advance). ``log_after_update`` plants a
    mistake for the tests that show the golden data can catch it: the
    quantities are then evaluated from the updated state."""
    from myokit_amd import _expressions as ex
    w, states, percell, order, names = model_c_parts(model, bound, fields)
    out = ['#include <math.h>', '#include "mk_math.h"',
           'const int oracle_n_state = %d;' % len(states),
           'const int oracle_n_field = %d;' % len(fields),
           'const int oracle_n_quant = %d;' % len(quantities),
           'void oracle_cell(double* y, double t, double dt, double pace,'
           ' double idiff, const double* fv, double* q, int log)',
           '{']
    out.append('    (void)t; (void)pace; (void)idiff; (void)fv;')
    for i in range(len(states)):
        out.append('    const double x%d = y[%d];' % (i, i))
    for i, var in enumerate(percell):
        out.append('    const double pc%d = %s;' % (i, w.w(var.rhs())))
    for var in order:
        if var.is_state():
            i = states.index(var)
            out.append('    const double dx%d = %s;' % (i, w.w(var.rhs())))
            out.append('    (void)dx%d;' % i)
        else:
            out.append('    const double %s = %s;'
                       % (names[ex.Name(var)], w.w(var.rhs())))
    log_block = ['    if (log) {']
    for j, var in enumerate(quantities):
        log_block.append('        q[%d] = %s;' % (j, w.w(ex.Name(var))))
    log_block.append('    }')
    if not log_after_update:
        out.extend(log_block)
    for i, s in enumerate(states):
        if s in rl_states:
            inf, tau = rl_states[s]
            inf, tau = w.w(ex.Name(inf)), w.w(ex.Name(tau))
            out.append('    y[%d] = %s - (%s - x%d) * mk_exp(-dt / %s);'
                       % (i, inf, inf, i, tau))
        else:
            out.append('    y[%d] = x%d + dt * dx%d;' % (i, i, i))
    if log_after_update:
        # the planted mistake: every logged name re-evaluated from the
        # updated state (a nested scope shadows the step's values)
        out.append('    {')
        for i in range(len(states)):
            out.append('    const double x%d = y[%d]; (void)x%d;' % (i, i, i))
        for var in order:
            if var.is_state():
                i = states.index(var)
                out.append('    const double dx%d = %s; (void)dx%d;'
                           % (i, w.w(var.rhs()), i))
            else:
                out.append('    const double %s = %s; (void)%s;'
                           % (names[ex.Name(var)], w.w(var.rhs()),
                              names[ex.Name(var)]))
        out.extend(log_block)
        out.append('    }')
    out.append('}')
    return '\n'.join(out) + '\n'


class OracleArgs(ctypes.Structure):
    _fields_ = [
        ('nx', ctypes.c_long), ('ny', ctypes.c_long),
        ('coupling', ctypes.c_int), ('i_vm', ctypes.c_int),
        ('gx', ctypes.c_double), ('gy', ctypes.c_double),
        ('gx_field', ctypes.c_void_p), ('gy_field', ctypes.c_void_p),
        ('n_conn', ctypes.c_long), ('conn_i', ctypes.c_void_p),
        ('conn_j', ctypes.c_void_p), ('conn_g', ctypes.c_void_p),
        ('pace_mode', ctypes.c_int),
        ('px0', ctypes.c_long), ('px1', ctypes.c_long),
        ('py0', ctypes.c_long), ('py1', ctypes.c_long),
        ('pace_mask', ctypes.c_void_p), ('fields', ctypes.c_void_p),
        ('state', ctypes.c_void_p),
        ('tmin', ctypes.c_double), ('tmax', ctypes.c_double),
        ('dt', ctypes.c_double), ('events', ctypes.c_void_p),
        ('n_events', ctypes.c_long), ('log_interval', ctypes.c_double),
        ('colmap', ctypes.c_void_p), ('n_cols', ctypes.c_int),
        ('log_values', ctypes.c_void_p), ('log_time', ctypes.c_void_p),
        ('log_pace', ctypes.c_void_p), ('max_rows', ctypes.c_long),
        ('nan_check', ctypes.c_int), ('threads', ctypes.c_int),
        ('rows', ctypes.c_long), ('t_end', ctypes.c_double),
        ('halted', ctypes.c_int), ('steps', ctypes.c_long),
    ]


def build(source):
    """Compiles a model source with the oracle driver; returns a CDLL."""
    os.makedirs(BUILD, exist_ok=True)
    with open(DRIVER_C) as f:
        drv = f.read()
    with open(MATH_H) as f:
        mh = f.read()
    h = hashlib.sha1((source + drv + mh).encode()).hexdigest()[:16]
    so = os.path.join(BUILD, 'oracle_' + h + '.so')
    if not os.path.isfile(so):
        src = os.path.join(BUILD, 'model_' + h + '.c')
        with open(src, 'w') as f:
            f.write(source)
        cmd = ['gcc', '-O2', '-fPIC', '-shared', '-ffp-contract=off',
               '-fno-fast-math', '-fopenmp'] + cpu_flags() + ['-I',
               os.path.dirname(MATH_H), '-o', so + '.tmp', src, DRIVER_C,
               '-lm']
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError('Oracle build failed:\n' + r.stderr)
        os.replace(so + '.tmp', so)
    lib = ctypes.CDLL(so)
    lib.oracle_run.restype = ctypes.c_int
    lib.oracle_run.argtypes = [ctypes.POINTER(OracleArgs)]
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class OracleSimulation:
    """Runs the oracle with a configuration like SimulationOpenCL's.

    ``model`` is an unprepared myokit_amd Model; preparation (RL
    conversion, bindings) follows the reference.
    """

    def __init__(self, model, protocol=None, ncells=256, diffusion=True,
                 rl=False):
        from myokit_amd.simulation import _prepare_model
        if isinstance(ncells, (tuple, list)):
            self.nx, self.ny = int(ncells[0]), int(ncells[1])
            self.dims = 2
        else:
            self.nx, self.ny = int(ncells), 1
            self.dims = 1
        self.n = self.nx * self.ny
        self.diffusion = bool(diffusion)
        self.model, self.vm, self.rl_states, self.bound = _prepare_model(
            model, rl, diffusion)
        self.protocol = protocol
        self.gx, self.gy = 10.0, 5.0
        self.gx_field = self.gy_field = None
        self.connections = None
        self.paced = (5, 1 if self.dims == 1 else 5, 0, 0) if diffusion \
            else None
        self.fields = {}
        self.dt = 0.005
        self.time = 0.0
        init = np.array(self.model.initial_values(True))
        self.state = np.tile(init, self.n)
        self.threads = int(os.environ.get('OMP_NUM_THREADS', '0')) or \
            len(os.sched_getaffinity(0))
        # planted mistake for the tests of the pins (see model_source)
        self.log_after_update = False

    def set_paced_cells(self, nx=5, ny=5, x=0, y=0):
        # Normalisation of SimulationOpenCL.set_paced_cells
        # (openclsim.py:1569-1593)
        nx, x = int(nx), int(x)
        if nx < 0:
            nx = -nx
            x -= nx
        if x < 0:
            x += self.nx
        if self.dims == 1:
            ny, y = 1, 0
        else:
            ny, y = int(ny), int(y)
            if ny < 0:
                ny = -ny
                y -= ny
            if y < 0:
                y += self.ny
        self.paced = (nx, ny, x, y)

    def set_paced_cell_list(self, cells):
        out = []
        for c in cells:
            if self.dims == 1:
                out.append(int(c))
            else:
                out.append(int(c[0]) + int(c[1]) * self.nx)
        self.paced = out

    def run(self, duration, log_vars=('membrane.V',), log_interval=1.0,
            cells=None, max_rows=None):
        """Runs ``duration``; returns a dict with ``time``, ``pace`` and,
        per logged variable, an array (rows, cells) of values."""
        m = self.model
        quantities = [m.get(v) for v in log_vars]
        fields = list(self.fields.keys())
        src = model_source(m, self.bound, self.rl_states,
                           [m.get(f) for f in fields], quantities,
                           log_after_update=self.log_after_update)
        lib = build(src)
        tmin = self.time
        tmax = tmin + duration
        if self.connections is not None:
            coupling = 3
        elif not self.diffusion:
            coupling = 0
        elif self.gx_field is not None:
            coupling = 2
        else:
            coupling = 1
        cells = np.arange(self.n) if cells is None else np.asarray(cells)
        nq = len(quantities)
        colmap = np.full((nq, self.n), -1, dtype=np.int32)
        for q in range(nq):
            colmap[q, cells] = q * len(cells) + np.arange(len(cells))
        ncols = nq * len(cells)
        if max_rows is None:
            max_rows = int(duration / log_interval) + 3
        vals = np.zeros((max_rows, max(ncols, 1)))
        lt = np.zeros(max_rows)
        lp = np.zeros(max_rows)
        a = OracleArgs()
        a.nx, a.ny = self.nx, self.ny
        a.coupling = coupling
        a.i_vm = self.vm.index() if self.vm is not None else 0
        a.gx, a.gy = self.gx, self.gy
        keep = []

        def arr(x, dtype):
            y = np.ascontiguousarray(x, dtype=dtype)
            keep.append(y)
            return _p(y)

        if self.gx_field is not None:
            a.gx_field = arr(self.gx_field, np.float64)
            a.gy_field = arr(self.gy_field if self.gy_field is not None
                             else [0.0], np.float64)
        if self.connections is not None:
            c = np.asarray(self.connections, dtype=np.float64).reshape(-1, 3)
            a.n_conn = len(c)
            a.conn_i = arr(c[:, 0], np.int64)
            a.conn_j = arr(c[:, 1], np.int64)
            a.conn_g = arr(c[:, 2], np.float64)
        if not self.diffusion:
            a.pace_mode = 0
        elif isinstance(self.paced, tuple):
            nx, ny, x, y = self.paced
            a.pace_mode = 1
            a.px0, a.px1, a.py0, a.py1 = x, x + nx, y, y + ny
        elif self.paced:
            a.pace_mode = 2
            mask = np.zeros(self.n, dtype=np.uint8)
            mask[self.paced] = 1
            a.pace_mask = arr(mask, np.uint8)
        else:
            a.pace_mode = 3
        if fields:
            fd = np.stack([self.fields[f] for f in fields], axis=1)
            a.fields = arr(fd.reshape(-1), np.float64)
        st = np.ascontiguousarray(self.state, dtype=np.float64)
        a.state = _p(st)
        a.tmin, a.tmax, a.dt = tmin, tmax, self.dt
        ev = np.zeros((0, 5)) if self.protocol is None else \
            self.protocol.to_arrays()
        a.events = arr(ev.reshape(-1) if len(ev) else [0.0], np.float64)
        a.n_events = len(ev)
        a.log_interval = 1e-9 if log_interval is None or log_interval <= 0 \
            else float(log_interval)
        a.colmap = arr(colmap.reshape(-1), np.int32)
        a.n_cols = ncols
        a.log_values = _p(vals)
        a.log_time = _p(lt)
        a.log_pace = _p(lp)
        a.max_rows = max_rows
        a.nan_check = int(any(q.is_state() for q in quantities))
        a.threads = self.threads
        r = lib.oracle_run(ctypes.byref(a))
        if r != 0:
            raise RuntimeError('Oracle pacing error ' + str(r))
        rows = min(a.rows, max_rows)
        self.state = st
        self.time = tmax
        out = {'time': lt[:rows].copy(), 'pace': lp[:rows].copy(),
               'rows': a.rows, 't_end': a.t_end, 'halted': bool(a.halted),
               'steps': a.steps}
        for q, v in enumerate(log_vars):
            out[v] = vals[:rows, q * len(cells):(q + 1) * len(cells)].copy()
        return out


class OracleFtArgs(ctypes.Structure):
    _fields_ = [
        ('cell_f', ctypes.c_void_p), ('cell_t', ctypes.c_void_p),
        ('nsf', ctypes.c_int), ('nst', ctypes.c_int),
        ('ivf', ctypes.c_int), ('ivt', ctypes.c_int),
        ('nfx', ctypes.c_long), ('nfy', ctypes.c_long),
        ('ntx', ctypes.c_long), ('nty', ctypes.c_long),
        ('ctx', ctypes.c_long), ('cty', ctypes.c_long),
        ('nx_paced', ctypes.c_long),
        ('gfx', ctypes.c_double), ('gfy', ctypes.c_double),
        ('gtx', ctypes.c_double), ('gty', ctypes.c_double),
        ('gft', ctypes.c_double),
        ('state_f', ctypes.c_void_p), ('state_t', ctypes.c_void_p),
        ('tmin', ctypes.c_double), ('tmax', ctypes.c_double),
        ('dt', ctypes.c_double), ('events', ctypes.c_void_p),
        ('n_events', ctypes.c_long), ('log_interval', ctypes.c_double),
        ('log_v', ctypes.c_void_p), ('log_time', ctypes.c_void_p),
        ('log_pace', ctypes.c_void_p), ('max_rows', ctypes.c_long),
        ('rows', ctypes.c_long), ('t_end', ctypes.c_double),
        ('steps', ctypes.c_long),
    ]


class OracleFiberTissue:
    """The oracle of FiberTissueSimulation (``oracle_ft_run``): double
    precision, forward Euler, arguments as the reference's constructor
    (``myokit/_sim/fiber_tissue.py:120-125``)."""

    def __init__(self, fiber_model, tissue_model, protocol=None,
                 ncells_fiber=(128, 2), ncells_tissue=(128, 128), nx_paced=5,
                 g_fiber=(9, 6), g_tissue=(9, 6), g_fiber_tissue=9,
                 dt=0.005):
        import re
        from myokit_amd.simulation import _prepare_model
        mf, self.vmf, _, bf = _prepare_model(fiber_model, False, True)
        mt, self.vmt, _, bt = _prepare_model(tissue_model, False, True)
        self.mf, self.mt = mf, mt
        src_f = model_source(mf, bf, {}, [], [])
        src_t = model_source(mt, bt, {}, [], [])
        for name in ('oracle_cell', 'oracle_n_state', 'oracle_n_field',
                     'oracle_n_quant'):
            src_t = re.sub(r'\b' + name + r'\b', name + '_t', src_t)
        self.lib = build(src_f + src_t)
        self.lib.oracle_ft_run.restype = ctypes.c_int
        self.lib.oracle_ft_run.argtypes = [ctypes.POINTER(OracleFtArgs)]
        self.nf = tuple(int(x) for x in ncells_fiber)
        self.nt = tuple(int(x) for x in ncells_tissue)
        self.nx_paced = min(int(nx_paced), self.nf[0])
        self.gf = tuple(float(x) for x in g_fiber)
        self.gt = tuple(float(x) for x in g_tissue)
        self.gft = float(g_fiber_tissue)
        self.cty = int(0.5 * (self.nt[1] - self.nf[1]))
        self.protocol = protocol
        self.dt = float(dt)
        self.time = 0.0
        self.state_f = np.tile(np.array(mf.initial_values(True)),
                               self.nf[0] * self.nf[1])
        self.state_t = np.tile(np.array(mt.initial_values(True)),
                               self.nt[0] * self.nt[1])

    def run(self, duration, log_interval=1.0):
        """Runs ``duration``; returns ``(time, pace, Vf[rows, nfx*nfy],
        Vt[rows, ntx*nty])``."""
        nf = self.nf[0] * self.nf[1]
        nt = self.nt[0] * self.nt[1]
        max_rows = int(duration / log_interval) + 3
        log_v = np.zeros((max_rows, nf + nt))
        lt = np.zeros(max_rows)
        lp = np.zeros(max_rows)
        events = np.ascontiguousarray(
            self.protocol.to_arrays() if self.protocol is not None
            else np.zeros((0, 5)), dtype=np.float64).reshape(-1)
        sf = np.ascontiguousarray(self.state_f, dtype=np.float64)
        st = np.ascontiguousarray(self.state_t, dtype=np.float64)
        a = OracleFtArgs()
        a.cell_f = ctypes.cast(self.lib.oracle_cell, ctypes.c_void_p).value
        a.cell_t = ctypes.cast(self.lib.oracle_cell_t, ctypes.c_void_p).value
        a.nsf = self.mf.count_states()
        a.nst = self.mt.count_states()
        a.ivf, a.ivt = self.vmf.index(), self.vmt.index()
        a.nfx, a.nfy = self.nf
        a.ntx, a.nty = self.nt
        a.ctx, a.cty = 0, self.cty
        a.nx_paced = self.nx_paced
        a.gfx, a.gfy = self.gf
        a.gtx, a.gty = self.gt
        a.gft = self.gft
        a.state_f, a.state_t = _p(sf), _p(st)
        a.tmin, a.tmax, a.dt = self.time, self.time + duration, self.dt
        a.events, a.n_events = _p(events), len(events) // 5
        a.log_interval = log_interval
        a.log_v, a.log_time, a.log_pace = _p(log_v), _p(lt), _p(lp)
        a.max_rows = max_rows
        r = self.lib.oracle_ft_run(ctypes.byref(a))
        if r:
            raise RuntimeError('oracle_ft_run failed: ' + str(r))
        self.state_f, self.state_t = sf, st
        self.time += duration
        n = min(a.rows, max_rows)
        return lt[:n], lp[:n], log_v[:n, :nf], log_v[:n, nf:]
