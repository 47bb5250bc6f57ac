"""
``SimulationOpenCL``: drop-in replacement for ``myokit.SimulationOpenCL``
(``myokit/_sim/openclsim.py:23``) that runs on an AMD Instinct GPU through
``libmyokit_amd.so``.

Same constructor ``(model, protocol=None, ncells=256, diffusion=True,
precision=SINGLE_PRECISION, native_maths=False, rl=False)``, same methods
(``run``, ``pre``, ``reset``, ``set_conductance``, ``set_conductance_field``,
``set_connections``, ``set_field``, ``set_paced_cells``,
``set_paced_cell_list``, ``set_step_size``, ``set_state``,
``set_default_state``, ``set_time``, ``set_protocol``, ``set_constant``,
``state``, ``default_state``, ``shape``, ``neighbors``, ``is_paced``,
``monodomain_conductance``, ``find_nan``), same argument checks and error
messages (``openclsim.py:149-322, 716-885, 1286-1757``), and the same
``DataLog`` keys. The time-stepping, logging and pacing rules are those of
the reference host module (``openclsim.c:1036-1211``), implemented in the
C++ library.
"""
import ctypes
import time as _time
from collections import OrderedDict

import numpy as np

from . import _datalog
from . import _expressions as ex
from . import _kernelgen
from . import _native
from ._datalog import (DOUBLE_PRECISION, LOG_BOUND, LOG_INTER, LOG_NONE,
                       LOG_STATE, SINGLE_PRECISION, dimco, float_str,
                       prepare_log, split_key)
from ._err import (FindNanError, NumericalError, SimulationCancelledError,
                   SimulationError, SimultaneousProtocolEventError)
from ._model import Variable


def _prepare_bindings(model, labels):
    """Maps bound variables to internal inputs; drops other bindings
    (behaviour of ``myokit._prepare_bindings``, ``_aux.py:434-475``)."""
    unused = []
    variables = {}
    for label, var in model.bindings():
        if label in labels:
            variables[var] = labels[label]
            var.set_rhs(0)
        else:
            unused.append(var)
    for var in unused:
        var.set_binding(None)
    return variables


def _prepare_model(model, rl, diffusion):
    """Clones the model (RL-converted if requested); returns
    ``(model, vm, rl_states, bound)``."""
    from .lib import hh
    vm = model.label('membrane_potential')
    rl_states = {}
    if rl:
        m = hh.convert_hh_states_to_inf_tau_form(model, vm)
        vm = m.get(vm.qname())
        for s in m.states():
            r = hh.get_inf_and_tau(s, vm)
            if r is not None:
                rl_states[s] = r
    else:
        m = model.clone()
        vm = None if vm is None else m.get(vm.qname())
    inputs = {'time': 'time', 'pace': 'pace'}
    if diffusion:
        inputs['diffusion_current'] = 'idiff'
    bound = _prepare_bindings(m, inputs)
    m.validate()
    m.create_unique_names()
    return m, vm, rl_states, bound


def _prepare_cell_model(model, rl, diffusion, fields, quantities):
    """Builds a CellModel from an unprepared model (used by build checks)."""
    m, vm, rl_states, bound = _prepare_model(model, rl, diffusion)
    fields = [m.get(f) if isinstance(f, str) else f for f in fields]
    quantities = [m.get(q) if isinstance(q, str) else q for q in quantities]
    return _kernelgen.CellModel(m, bound, rl_states, fields, quantities)


# Compiled programs, keyed by kernel source and options
_PROGRAMS = {}


def _program(source, options, device=0):
    key = (source, options, device)
    p = _PROGRAMS.get(key)
    if p is None:
        import os
        cache = os.environ.get('MYOKIT_AMD_CACHE')
        if cache:
            os.makedirs(cache, exist_ok=True)
        prog = ctypes.c_void_p()
        log = ctypes.create_string_buffer(1 << 16)
        _native.check(_native.lib().mkh_program_build(
            source.encode(), options.encode(),
            cache.encode() if cache else None, device, ctypes.byref(prog),
            log, 1 << 16), 'mkh_program_build')
        p = _PROGRAMS[key] = prog
        _native.track_program(prog)
    return p


def _graph_copies(i, j, n, block_cells, min_copies=4):
    """The number K of identical copies a graph of ``n`` cells with
    connections ``(i, j)`` is made of: the largest K >= ``min_copies``
    (dividing ``n``) such that no connection crosses a block of n / K
    cells and every block holds the same connections as the first, shifted
    by its offset (in any order); 1 if there is none, or if a copy has at
    most ``block_cells`` cells or is a set of chains (the block and chain
    drivers, which need the copies contiguous, then apply)."""
    if len(i) == 0:
        return 1
    lo, hi = np.minimum(i, j), np.maximum(i, j)
    divisors = set()
    d = 1
    while d * d <= n:
        if n % d == 0:
            divisors.update((d, n // d))
        d += 1
    for k in sorted(divisors, reverse=True):
        if k < min_copies:
            break
        m = n // k
        if m <= block_cells:
            continue
        b = lo // m
        if not np.array_equal(b, hi // m):
            continue
        counts = np.bincount(b, minlength=k)
        if not np.all(counts == counts[0]):
            continue
        key = (lo - b * m) * m + (hi - b * m)
        order = np.lexsort((key, b))
        keys = key[order].reshape(k, -1)
        if not np.all(keys == keys[0]):
            continue
        first = b == 0
        deg = np.bincount(np.concatenate([lo[first], hi[first]]),
                          minlength=m)
        if deg.max() <= 2:
            return 1
        return k
    return 1


class _Engine:
    """One device simulation handle for one compiled kernel."""

    def __init__(self, program, config):
        self.config = config
        h = ctypes.c_void_p()
        _native.check(_native.lib().mkh_sim_create(
            program, ctypes.byref(config), ctypes.byref(h)),
            'mkh_sim_create')
        self.h = h
        _native.track(self, _native.lib().mkh_sim_destroy)

    def __del__(self):
        try:
            if self.h and not _native.closed:
                _native.untrack(self)
                _native.lib().mkh_sim_destroy(self.h)
        except Exception:   # pragma: no cover
            pass
        self.h = None


class SimulationOpenCL:
    """Multi-cell (1D, 2D or graph) simulation of a cell model on a GPU.

    See ``myokit.SimulationOpenCL`` for the model of use: forward Euler (or
    Rush-Larsen for Hodgkin-Huxley gates when ``rl=True``) with a fixed step
    size, diffusion current ``i = sum g * (V - V_j)`` over neighbours, cells
    paced by a protocol, and per-cell logging.
    """

    # The current state lives on the device between runs. The host copy is
    # refreshed on access (``_state`` getter) and uploaded only after the
    # host changed it.
    @property
    def _state(self):
        if self._dev_newer:
            self._download_state()
        return self._state_arr

    @_state.setter
    def _state(self, value):
        self._state_arr = np.array(value, dtype=np.float64)
        self._dev_newer = False
        self._host_dirty = True

    def _download_state(self):
        st = np.zeros(self._ntotal * self._nstate)
        _native.check(_native.lib().mkh_sim_get_state(
            self._engine.h, _native.ptr(st, ctypes.c_double), len(st)),
            'get_state')
        perm, iperm = self._cell_order()
        if iperm is not None:
            st = st.reshape(self._ntotal, self._nstate)[iperm].reshape(-1)
        self._state_arr = st
        self._dev_newer = False

    def _max_block_cells(self):
        """Cells of the largest domain the one-workgroup-per-domain driver
        holds in LDS (MK_BCELLS)."""
        return 2048 if self._precision == DOUBLE_PRECISION else 4096

    def _cell_order(self):
        """``(perm, iperm)``: the order in which the device holds the cells
        of a set_connections graph (device cell k is the user's cell
        perm[k], user cell c is device cell iperm[c]), or ``(None, None)``
        for the user's own order. Graphs are held in reverse Cuthill-McKee
        order, which puts neighbours close together in memory: the
        diffusion current's and the adjoint's gathers of neighbour values
        then touch few cache lines (a random geometric graph of 262,144
        cells numbered at random has neighbours 87,000 indices apart on
        average, 330 after reordering). The per-cell arithmetic, the order
        of each cell's connections included, is the same in every order.

        A graph that is the disjoint union of identical copies (a batch of
        parameter sets run as one simulation, each set a copy of the same
        graph in its own block of cells) is held with the copies
        interleaved, cell p of every copy side by side, each copy in the
        reverse Cuthill-McKee order of the first: a wave then reads each
        neighbour's potential for many copies in one contiguous segment.
        Copies small enough for one workgroup each keep the block order,
        which the one-workgroup-per-domain driver needs."""
        if self._connections is None or len(self._connections) == 0:
            return None, None
        c = self._order_cache
        if c is not None and c[0] == self._graph_version:
            return c[1], c[2]
        from scipy.sparse import coo_matrix
        from scipy.sparse.csgraph import reverse_cuthill_mckee
        a = np.asarray(self._connections, dtype=np.float64).reshape(-1, 3)
        n = self._ntotal
        i = a[:, 0].astype(np.int64)
        j = a[:, 1].astype(np.int64)
        copies = _graph_copies(i, j, n, self._max_block_cells())
        m = n // copies
        if copies > 1:
            # one copy's order, the copies interleaved: device cell
            # p * copies + k is cell p of copy k
            i, j = i[i < m], j[i < m]
        g = coo_matrix((np.ones(2 * len(i), dtype=np.int8),
                        (np.concatenate([i, j]), np.concatenate([j, i]))),
                       shape=(m, m)).tocsr()
        perm = np.asarray(reverse_cuthill_mckee(g, symmetric_mode=True),
                          dtype=np.int64)
        if copies > 1:
            perm = (perm[:, None] + m * np.arange(copies)[None, :]).ravel()
        iperm = np.empty(n, dtype=np.int64)
        iperm[perm] = np.arange(n)
        if np.array_equal(perm, np.arange(n)):
            perm = iperm = None
        self._order_cache = (self._graph_version, perm, iperm)
        self._copies = copies
        self._pcolmaps = {}
        return perm, iperm

    def _before_reorder(self):
        """Called before the device's cell order may change (a new graph,
        or a graph replaced by a stencil): a newer state on the device comes
        back in the old order, and goes up again in the new one."""
        if self._dev_newer and self._engine is not None:
            self._download_state()
        self._host_dirty = True

    def _touch(self, what='config'):
        # Marks device-side configuration as changed
        if what == 'graph':
            self._graph_version += 1
        self._config_version += 1

    def __init__(self, model, protocol=None, ncells=256, diffusion=True,
                 precision=SINGLE_PRECISION, native_maths=False, rl=False):
        self._config_version = 0
        self._graph_version = 0
        self._dev_newer = False
        self._engine = None
        self._engines = OrderedDict()   # kernel key -> engine (recent last)
        # generator inputs -> (kernel source, CellModel): the source of a
        # repeated run is not generated again
        self._sources = {}
        # number of quantities -> column map of run_block (every cell)
        self._block_colmaps = {}
        # device cell order of a set_connections graph (_cell_order) and
        # column maps permuted to it
        self._order_cache = None
        self._pcolmaps = {}
        if not model.is_valid():
            model.validate()
        if model.has_interdependent_components():
            cycles = '\n'.join(
                '  ' + ' > '.join(x.name() for x in c)
                for c in model.component_cycles())
            raise ValueError(
                'This simulation requires models without interdependent'
                ' components. Please restructure the model and re-run.'
                '\nCycles:\n' + cycles)
        self.set_protocol(protocol)
        try:
            self._nx = int(ncells)
            self._ny = 1
            self._dims = (self._nx,)
        except TypeError:
            try:
                assert len(ncells) == 2
            except (TypeError, AssertionError):
                raise ValueError(
                    'The argument "ncells" must be either a scalar or a tuple'
                    ' (nx, ny).')
            self._nx = int(ncells[0])
            self._ny = int(ncells[1])
            self._dims = (self._nx, self._ny)
        if self._nx < 1 or self._ny < 1:
            raise ValueError(
                'The number of cells in any direction must be at least 1.')
        self._ntotal = self._nx * self._ny
        self._diffusion_enabled = bool(diffusion)
        if precision not in (SINGLE_PRECISION, DOUBLE_PRECISION):
            raise ValueError('Only single and double precision are supported.')
        self._precision = precision
        self._native_math = bool(native_maths)
        self._rl = bool(rl)

        vm = model.label('membrane_potential')
        if self._diffusion_enabled or self._rl:
            if vm is None:
                raise ValueError(
                    'This simulation requires the membrane potential'
                    ' variable to be labelled as "membrane_potential".')
            if not vm.is_state():
                raise ValueError(
                    'The variable labelled as membrane potential must'
                    ' be a state variable.')

        self._model, self._vm, self._rl_states, self._bound_variables = \
            _prepare_model(model, self._rl, self._diffusion_enabled)

        self._gx = self._gy = None
        if self._diffusion_enabled:
            self.set_conductance()
        self._gx_field = None
        self._gy_field = None
        self._connections = None
        self._paced_cells = None
        if self._diffusion_enabled:
            self.set_paced_cells()
        self._fields = OrderedDict()
        self.set_step_size()
        self._time = 0
        self._nstate = self._model.count_states()
        init = np.array(self._model.initial_values(True), dtype=np.float64)
        self._state = np.tile(init, self._ntotal)
        self._default_state = self._state.copy()
        self._global = ['time', 'pace']
        self._engine_key = None
        self._device = 0
        # threads per workgroup; None: chosen per layout (_threads)
        self._block_threads = None
        self._domains_cache = None
        self._copies = 1
        self._tuning = {}
        self._options = {}
        self.last_run_stats = {}

    # ------------------------------------------------------------------
    # Simple accessors
    # ------------------------------------------------------------------
    def conductance(self):
        if not self._diffusion_enabled:
            raise RuntimeError(
                'This method is unavailable when diffusion is disabled.')
        if (self._gx_field is not None) or (self._connections is not None):
            return None
        if len(self._dims) == 1:
            return self._gx
        return (self._gx, self._gy)

    def _cell_slice(self, x, y, src):
        x = int(x or 0)
        if x < 0 or x >= self._nx:
            raise IndexError('Given x-index out of range.')
        y = int(y or 0)
        if len(self._dims) == 2:
            if y < 0 or y >= self._ny:
                raise IndexError('Given y-index out of range.')
            x += y * self._nx
        elif y != 0:
            raise ValueError(
                'Y-coordinate specified for 1-dimensional simulation.')
        return list(src[x * self._nstate:(x + 1) * self._nstate])

    def default_state(self, x=None, y=None):
        if x is None:
            return list(self._default_state)
        return self._cell_slice(x, y, self._default_state)

    def state(self, x=None, y=None):
        if x is None:
            return list(self._state)
        return self._cell_slice(x, y, self._state)

    def is_2d(self):
        return len(self._dims) == 2

    def is_paced(self, x, y=None):
        if not self._diffusion_enabled:
            raise RuntimeError(
                'This method is unavailable when diffusion is disabled.')
        x = int(x)
        if x < 0 or x >= self._dims[0]:
            raise IndexError('X-coordinate out of range: ' + str(x) + '.')
        if len(self._dims) == 2:
            if y is None:
                raise ValueError(
                    'No y-coordinate specified in 2-dimensional simulation.')
            y = int(y)
            if y < 0 or y >= self._dims[1]:
                raise IndexError('Y-coordinate out of range: ' + str(y) + '.')
        elif y is None:
            y = 0
        elif y != 0:
            raise ValueError(
                'Y-coordinate specified for 1-dimensional simulation.')
        if type(self._paced_cells) == tuple:
            nx, ny, xmin, ymin = self._paced_cells
            return (xmin <= x < xmin + nx and ymin <= y < ymin + ny)
        return (x + y * self._dims[0]) in self._paced_cells

    def monodomain_conductance(self, chi, k, D, dx, A=1):
        """``g = (1 / chi) * (k / (k + 1)) * D * A / dx^2``
        (``openclsim.py:753-812``)."""
        return D * A * k / ((k + 1) * chi * dx * dx)

    def neighbors(self, x, y=None):
        if not self._diffusion_enabled:
            raise RuntimeError(
                'This method is unavailable when diffusion is disabled.')
        x = int(x or 0)
        if x < 0 or x >= self._dims[0]:
            raise IndexError('X-coordinate out of range: ' + str(x) + '.')
        if len(self._dims) == 2:
            if y is None:
                raise ValueError(
                    'No y-coordinate specified in 2-dimensional simulation.')
            y = int(y)
            if y < 0 or y >= self._dims[1]:
                raise IndexError('Y-coordinate out of range: ' + str(y) + '.')
        elif not (y is None or y == 0):
            raise ValueError(
                'Y-coordinate specified for 1-dimensional simulation.')
        if self._connections is not None:
            out = []
            for i, j, c in self._connections:
                if i == x:
                    out.append(int(j))
                elif j == x:
                    out.append(int(i))
            return out
        out = []
        if x > 0:
            out.append(x - 1)
        if x + 1 < self._dims[0]:
            out.append(x + 1)
        if len(self._dims) == 2:
            out = [(i, y) for i in out]
            if y > 0:
                out.append((x, y - 1))
            if y + 1 < self._dims[1]:
                out.append((x, y + 1))
        return out

    def shape(self):
        if len(self._dims) == 2:
            return (self._ny, self._nx)
        return self._nx

    def step_size(self):
        return self._step_size

    def time(self):
        return self._time

    # ------------------------------------------------------------------
    # Setters (openclsim.py:1286-1757)
    # ------------------------------------------------------------------
    def set_conductance(self, gx=10, gy=5):
        self._touch('config')
        if not self._diffusion_enabled:
            raise RuntimeError(
                'This method is unavailable when diffusion is disabled.')
        gx, gy = float(gx), float(gy)
        if gx < 0:
            raise ValueError('Invalid conductance gx=' + str(gx))
        if gy < 0:
            raise ValueError('Invalid conductance gy=' + str(gy))
        self._gx, self._gy = gx, gy
        self._gx_field = self._gy_field = None
        if getattr(self, '_connections', None) is not None:
            self._before_reorder()
            self._graph_version += 1
        self._connections = None

    def set_conductance_field(self, gx, gy=None):
        self._touch('config')
        if not self._diffusion_enabled:
            raise RuntimeError(
                'This method is unavailable when diffusion is disabled.')
        gx = np.asarray(gx, dtype=float)
        if len(self._dims) == 1:
            s = self._nx - 1
            if gx.shape != (s, ):
                raise ValueError(
                    'The argument `gx` must have length ' + str(s) + '.')
            if gy is not None:
                raise ValueError(
                    'For 1-d simulations the argument `gy` must be None.')
        else:
            s = (self._ny, self._nx - 1)
            if gx.shape != s:
                raise ValueError(
                    'The argument `gx` must have dimensions ' + str(s) + '.')
            if gy is None:
                raise ValueError(
                    'The argument `gy` must be set for 2-d simulations.')
            gy = np.asarray(gy, dtype=float)
            s = (self._ny - 1, self._nx)
            if gy.shape != s:
                raise ValueError(
                    'The argument `gy` must have dimensions ' + str(s) + '.')
        if np.any(gx < 0):
            raise ValueError(
                'The argument `gx` can not contain negative values.')
        if (gy is not None) and np.any(gy < 0):
            raise ValueError(
                'The argument `gy` can not contain negative values.')
        self._gx_field = list(gx.reshape((self._nx - 1) * self._ny))
        if gy is None:
            self._gy_field = None
        else:
            self._gy_field = list(gy.reshape(self._nx * (self._ny - 1)))
        if getattr(self, '_connections', None) is not None:
            self._before_reorder()
            self._graph_version += 1
        self._connections = None

    def set_connections(self, connections):
        if self._diffusion_enabled and len(self._dims) == 1 and \
                connections is not None:
            self._before_reorder()
        self._touch('graph')
        if not self._diffusion_enabled:
            raise RuntimeError(
                'This method is unavailable when diffusion is disabled.')
        if len(self._dims) != 1:
            raise RuntimeError('Connections can only be specified in 1d mode.')
        if connections is None:
            raise ValueError(
                'Connection list cannot be None. To unset connections, call '
                ' set_conductance() with the new conductance value.')
        if isinstance(connections, np.ndarray):
            self._connections = self._connection_array(connections)
            self._gx_field = self._gy_field = None
            return
        conns = []
        doubles = set()
        for x in connections:
            try:
                i, j, c = x
            except Exception:
                raise ValueError(
                    'Connections must be a list of 3-tuples'
                    ' (cell_index_1, cell_index_2, conductance).')
            i, j = int(i), int(j)
            if i == j or i < 0 or j < 0 or i >= self._nx or j >= self._nx:
                raise ValueError(
                    'Invalid connection: (' + str(i) + ', ' + str(j) + ', '
                    + str(c) + ')')
            i, j = (i, j) if i < j else (j, i)
            if (i, j) in doubles:
                raise ValueError(
                    'Duplicate connection: (' + str(i) + ', ' + str(j) + ', '
                    + str(c) + ')')
            doubles.add((i, j))
            c = float(c)
            if c < 0:
                raise ValueError('Invalid conductance: ' + str(c))
            conns.append((i, j, c))
        self._connections = conns
        self._gx_field = self._gy_field = None

    def _connection_array(self, a):
        """set_connections for an (E, 3) array of (cell, cell, conductance)
        rows: the checks and messages of the list form, vectorised for
        graphs of millions of connections. Returns the rows with the lower
        index first, as float64."""
        a = np.asarray(a, dtype=np.float64)
        if a.ndim != 2 or a.shape[1] != 3:
            raise ValueError(
                'Connections must be a list of 3-tuples'
                ' (cell_index_1, cell_index_2, conductance).')
        i = a[:, 0].astype(np.int64)
        j = a[:, 1].astype(np.int64)
        c = a[:, 2]
        bad = (i == j) | (i < 0) | (j < 0) | (i >= self._nx) | (j >= self._nx)
        if bad.any():
            k = int(np.argmax(bad))
            raise ValueError(
                'Invalid connection: (' + str(int(i[k])) + ', '
                + str(int(j[k])) + ', ' + str(c[k]) + ')')
        lo, hi = np.minimum(i, j), np.maximum(i, j)
        key = lo * self._nx + hi
        order = np.argsort(key, kind='stable')
        dup = np.nonzero(key[order][1:] == key[order][:-1])[0]
        if len(dup):
            k = int(order[dup[0] + 1])
            raise ValueError(
                'Duplicate connection: (' + str(int(lo[k])) + ', '
                + str(int(hi[k])) + ', ' + str(c[k]) + ')')
        if (c < 0).any():
            k = int(np.argmax(c < 0))
            raise ValueError('Invalid conductance: ' + str(c[k]))
        return np.stack([lo.astype(np.float64), hi.astype(np.float64), c],
                        axis=1)

    def set_constant(self, var, value):
        value = float(value)
        if isinstance(var, Variable):
            var = var.qname()
        var = self._model.get(var)
        if not var.is_literal():
            raise ValueError(
                'The given variable <' + var.qname() + '> is not a literal.')
        self._model.set_value(var.qname(), value)
        # the generated kernel inlines constants
        self._sources.clear()

    def set_default_state(self, state, x=None, y=None):
        self._default_state = self._set_state(state, x, y,
                                              self._default_state)

    def set_field(self, var, values):
        self._touch('config')
        if isinstance(var, Variable):
            var = var.qname()
        var = self._model.get(var)
        if var.is_bound():
            raise ValueError('Bound values cannot be replaced by fields.')
        if not var.is_constant():
            raise ValueError('Only constants can be used for fields.')
        values = np.asarray(values, dtype=float)
        if len(self._dims) == 1:
            if values.shape != (self._nx, ):
                raise ValueError(
                    'The argument `values` must have length ' + str(self._nx)
                    + '.')
        else:
            shape = (self._ny, self._nx)
            if values.shape != shape:
                raise ValueError(
                    'The argument `values` must have dimensions ' + str(shape)
                    + '.')
        self._fields[var] = values.reshape(self._nx * self._ny).copy()

    def remove_field(self, var):
        self._touch('config')
        if isinstance(var, Variable):
            var = var.qname()
        var = self._model.get(var)
        del self._fields[var]

    def set_paced_cells(self, nx=5, ny=5, x=0, y=0):
        self._touch('config')
        if not self._diffusion_enabled:
            raise RuntimeError(
                'This method is unavailable when diffusion is disabled.')
        nx = int(nx)
        x = int(x)
        if nx < 0:
            nx = -nx
            x -= nx
        if x < 0:
            x += self._nx
        if len(self._dims) == 1:
            ny = 1
            y = 0
        else:
            ny = int(ny)
            if ny < 0:
                ny = -ny
                y -= ny
            if y < 0:
                y += self._ny
        self._paced_cells = (nx, ny, x, y)

    def set_paced_cell_list(self, cells):
        self._touch('config')
        if not self._diffusion_enabled:
            raise RuntimeError(
                'This method is unavailable when diffusion is disabled.')
        paced = []
        if len(self._dims) == 1:
            for cell in cells:
                cell = int(cell)
                if cell < 0 or cell >= self._nx:
                    raise IndexError(
                        'Cell index out of range: ' + str(cell) + '.')
                paced.append(cell)
        else:
            for i, j in cells:
                i, j = int(i), int(j)
                if i < 0 or j < 0 or i >= self._nx or j >= self._ny:
                    raise IndexError(
                        'Cell index out of range: (' + str(i) + ', ' + str(j)
                        + ').')
                paced.append(i + j * self._nx)
        self._paced_cells = paced

    def set_protocol(self, protocol=None):
        self._protocol = None if protocol is None else protocol.clone()

    def _set_state(self, state, x, y, update):
        n = len(state)
        if n == self._nstate * self._ntotal:
            if x is not None:
                raise ValueError(
                    'A state for all cells was passed in, but argument x'
                    ' was not None.')
            if y is not None:
                raise ValueError(
                    'A state for all cells was passed in, but argument y'
                    ' was not None.')
            return np.array(state, dtype=np.float64)
        elif n != self._nstate:
            raise ValueError(
                'Given state must have the same size as a single'
                ' cell state or a full simulation state')
        if x is None:
            return np.tile(np.asarray(state, dtype=np.float64), self._ntotal)
        x = int(x or 0)
        if x < 0 or x >= self._nx:
            raise IndexError('Given x-index out of range.')
        y = int(y or 0)
        if len(self._dims) == 2:
            if y < 0 or y >= self._ny:
                raise IndexError('Given y-index out of range.')
            x += y * self._nx
        elif y != 0:
            raise ValueError('Y-index given for a 1-dimensional simulation.')
        update = np.array(update, dtype=np.float64)
        off = x * self._nstate
        update[off:off + self._nstate] = state
        return update

    def set_state(self, state, x=None, y=None):
        self._state = self._set_state(state, x, y, self._state)

    def set_step_size(self, step_size=0.005):
        step_size = float(step_size)
        if step_size <= 0:
            raise ValueError('Step size must be greater than zero.')
        self._step_size = step_size

    def set_time(self, time=0):
        self._time = float(time)

    # ------------------------------------------------------------------
    # Running
    # ------------------------------------------------------------------
    def pre(self, duration, report_nan=True, progress=None,
            msg='Pre-pacing SimulationOpenCL'):
        self._run(duration, LOG_NONE, 1, report_nan, progress, msg)
        self._default_state = self._state.copy()

    def reset(self):
        self._time = 0
        self._state = self._default_state.copy()

    def run(self, duration, log=None, log_interval=1.0, report_nan=True,
            progress=None, msg='Running SimulationOpenCL'):
        r = self._run(duration, log, log_interval, report_nan, progress, msg)
        self._time += duration
        return r

    def _coupling(self):
        if not self._diffusion_enabled:
            return _native.COUPLING_NONE
        if self._connections is not None:
            return _native.COUPLING_GRAPH
        if self._gx_field is not None:
            return _native.COUPLING_HETERO
        return _native.COUPLING_STENCIL

    def set_tuning(self, driver=None, block_threads=None, sheet_cpt=None,
                   sheet_waves=None, adjoint_waves=None, chain_cpt=None,
                   chain_waves=None, adjoint_cpt=None, grid_waves=None):
        """Launch configuration of the device engine (chosen automatically
        by default; :func:`myokit_amd.tune.autotune` measures and picks
        one). ``driver``: ``'auto'``, ``'cells'``, ``'block'``, ``'grid'``,
        ``'sheet'`` or ``'chain'``; ``block_threads``: threads per
        workgroup of the cells, block and grid kernels (multiple of 64);
        ``sheet_cpt``: cells per thread of the persistent sheet (tile
        64 x 4*cpt); ``sheet_waves``: waves per SIMD its register budget
        allows; ``chain_cpt`` and ``chain_waves``: the same for the chain
        driver (graphs of disjoint chains; one-wave tiles of at most
        64*cpt cells); ``adjoint_waves``: the waves per SIMD of the reverse
        kernels of :class:`TissueAdjoint` (0: as many as their registers
        allow); ``adjoint_cpt``: cells per thread of the persistent reverse
        sheet (tile 64 x 4*cpt); ``grid_waves``: the waves per SIMD the grid
        driver's register budget allows (0: no limit). Arguments left as
        None keep their current value."""
        names = {'auto': -1, 'cells': _native.DRIVER_CELLS,
                 'block': _native.DRIVER_BLOCK, 'grid': _native.DRIVER_GRID,
                 'sheet': _native.DRIVER_SHEET,
                 'chain': _native.DRIVER_CHAIN}
        if block_threads is not None:
            block_threads = int(block_threads)
            if block_threads % 64 or not 64 <= block_threads <= 1024:
                raise ValueError('block_threads must be a multiple of 64 in'
                                 ' [64, 1024].')
            self._block_threads = block_threads
        if sheet_cpt is not None:
            if int(sheet_cpt) not in _kernelgen.SHEET_CPT_CHOICES:
                raise ValueError('sheet_cpt must be one of '
                                 + str(_kernelgen.SHEET_CPT_CHOICES) + '.')
            self._tuning['sheet_cpt'] = int(sheet_cpt)
        if sheet_waves is not None:
            if not 1 <= int(sheet_waves) <= 8:
                raise ValueError('sheet_waves must be in [1, 8].')
            self._tuning['sheet_waves'] = int(sheet_waves)
        if chain_cpt is not None:
            if int(chain_cpt) not in _kernelgen.CHAIN_CPT_CHOICES:
                raise ValueError('chain_cpt must be one of '
                                 + str(_kernelgen.CHAIN_CPT_CHOICES) + '.')
            self._tuning['chain_cpt'] = int(chain_cpt)
        if chain_waves is not None:
            if not 1 <= int(chain_waves) <= 8:
                raise ValueError('chain_waves must be in [1, 8].')
            self._tuning['chain_waves'] = int(chain_waves)
        if grid_waves is not None:
            if not 0 <= int(grid_waves) <= 8:
                raise ValueError('grid_waves must be in [0, 8].')
            self._tuning['grid_waves'] = int(grid_waves)
        if adjoint_cpt is not None:
            if int(adjoint_cpt) not in _kernelgen.ADJ_SHEET_CPT_CHOICES:
                raise ValueError('adjoint_cpt must be one of '
                                 + str(_kernelgen.ADJ_SHEET_CPT_CHOICES)
                                 + '.')
            self._tuning['adjoint_cpt'] = int(adjoint_cpt)
        if adjoint_waves is not None:
            if not 0 <= int(adjoint_waves) <= 8:
                raise ValueError('adjoint_waves must be in [0, 8].')
            self._tuning['adjoint_waves'] = int(adjoint_waves)
        if driver is not None:
            if driver not in names:
                raise ValueError('Unknown driver ' + repr(driver) + '.')
            self._tuning['driver'] = driver
            self.set_engine_option('driver', names[driver])

    def _graph_domains(self):
        """(largest, count) of the connected components of the
        set_connections graph (isolated cells count as components of one),
        cached per graph."""
        c = self._domains_cache
        if c is not None and c[0] == self._graph_version:
            return c[1]
        from scipy.sparse import coo_matrix
        from scipy.sparse.csgraph import connected_components
        a = np.asarray(self._connections, dtype=np.float64).reshape(-1, 3)
        n = self._ntotal
        g = coo_matrix((np.ones(len(a)), (a[:, 0].astype(np.int64),
                                          a[:, 1].astype(np.int64))),
                       shape=(n, n))
        k, labels = connected_components(g, directed=False)
        r = (int(np.bincount(labels).max()), int(k))
        self._domains_cache = (self._graph_version, r)
        return r

    def _threads(self):
        """Threads per workgroup of the cells, block and grid kernels: the
        value given to :meth:`set_tuning`, else 256, except where the
        library will run the block layout (one workgroup per domain:
        graphs whose components all fit a workgroup's LDS tile, and single
        cables or sheets of at most 512 cells). There it is one thread per
        cell of the largest domain (rounded up to a multiple of 64, at most
        1024): one workgroup per cable leaves a GPU with a few waves per
        CU, and more threads hide more latency (256 cables of 50-2000
        cells: 3.99e10 cell-steps/s with 1024 threads, 2.05e10 with
        256)."""
        if self._block_threads is not None:
            return self._block_threads
        c = self._coupling()
        if c == _native.COUPLING_NONE:
            return 256
        cap = 2048 if self._precision == DOUBLE_PRECISION else 4096
        if c == _native.COUPLING_GRAPH:
            largest, _ = self._graph_domains()
            block = largest <= cap
        else:
            largest = self._ntotal
            block = largest <= min(cap, 512)
        if not block:
            return 256
        return int(min(1024, max(64, -(-largest // 64) * 64)))

    def tuning(self):
        """The current launch configuration (see :meth:`set_tuning`)."""
        return {'driver': self._tuning.get('driver', 'auto'),
                'block_threads': self._threads(),
                'sheet_cpt': self._tuning.get(
                    'sheet_cpt', _kernelgen.sheet_cpt(self._precision)),
                'sheet_waves': self._tuning.get(
                    'sheet_waves', _kernelgen.sheet_waves(self._precision)),
                'chain_cpt': self._tuning.get(
                    'chain_cpt', _kernelgen.chain_cpt(self._precision)),
                'chain_waves': self._tuning.get(
                    'chain_waves', _kernelgen.chain_waves(self._precision)),
                'adjoint_waves': self._tuning.get('adjoint_waves'),
                'adjoint_cpt': self._tuning.get(
                    'adjoint_cpt',
                    _kernelgen.adjoint_sheet_cpt(self._precision)),
                'grid_waves': self._tuning.get(
                    'grid_waves', _kernelgen.grid_waves(self._precision))}

    def set_engine_option(self, key, value):
        """Tuning option passed to the device library (``batch_steps``,
        ``log_rows``, ``time_kernels``, ``driver``)."""
        self._options[key] = int(value)
        if self._engine is not None:
            _native.check(_native.lib().mkh_sim_set_option(
                self._engine.h, key.encode(), int(value)), 'set_option')
            self._engine.options[key] = int(value)

    def engine_driver(self):
        """Name of the device execution driver of the last run (``cells``,
        ``block``, ``grid``, ``sheet`` or ``chain``), or None before the
        first run."""
        if self._engine is None:
            return None
        d = ctypes.c_int(-1)
        _native.check(_native.lib().mkh_sim_driver(self._engine.h,
                                                   ctypes.byref(d)), 'driver')
        return {0: 'cells', 1: 'block', 2: 'grid', 3: 'sheet',
                4: 'chain'}[d.value]

    def engine_timing(self, reset=False):
        """Kernel timing of the device engine (needs the ``time_kernels``
        option): ``{'kernel_ms', 'launches', 'cell_steps'}``."""
        if self._engine is None:
            return {'kernel_ms': 0.0, 'launches': 0, 'cell_steps': 0}
        ms = ctypes.c_double(0)
        nl = ctypes.c_longlong(0)
        cs = ctypes.c_longlong(0)
        L = _native.lib()
        _native.check(L.mkh_sim_get_timing(
            self._engine.h, ctypes.byref(ms), ctypes.byref(nl),
            ctypes.byref(cs)), 'get_timing')
        if reset:
            L.mkh_sim_reset_timing(self._engine.h)
        return {'kernel_ms': ms.value, 'launches': nl.value,
                'cell_steps': cs.value}

    def _kernel_source(self, quantities, adjoint=False,
                       adjoint_precision=None):
        """The generated kernel source and its CellModel, cached per
        simulation by everything the generator reads (the model's values
        too: ``set_constant`` clears the cache)."""
        fields = list(self._fields.keys())
        coupling = self._coupling()
        tune = self._tuning
        threads = self._threads()
        gkey = (tuple(fields), tuple(quantities), int(coupling),
                threads, self._native_math, bool(adjoint),
                adjoint_precision, tuple(sorted(tune.items())))
        hit = self._sources.get(gkey)
        if hit is not None:
            return hit
        cm = _kernelgen.CellModel(self._model, self._bound_variables,
                                  self._rl_states, fields, quantities)
        src = _kernelgen.generate_kernel(
            cm, self._precision, coupling, threads,
            self._native_math, adjoint=adjoint,
            adjoint_precision=adjoint_precision,
            tile_cpt=tune.get('sheet_cpt'),
            tile_waves=tune.get('sheet_waves'),
            adjoint_waves=tune.get('adjoint_waves'),
            chain_cells=tune.get('chain_cpt'),
            chain_budget=tune.get('chain_waves'),
            adjoint_cpt=tune.get('adjoint_cpt'),
            grid_budget=tune.get('grid_waves'))
        if len(self._sources) >= 16:
            self._sources.pop(next(iter(self._sources)))
        self._sources[gkey] = (src, cm)
        return src, cm

    def _get_engine(self, quantities, adjoint=False, adjoint_precision=None):
        src, cm = self._kernel_source(quantities, adjoint, adjoint_precision)
        fields = list(self._fields.keys())
        coupling = self._coupling()

        opts = _kernelgen.compile_options(self._precision)
        key = (src, opts, self._ntotal, self._dims)
        if self._engine is not None and self._engine_key == key:
            return self._engine, cm
        eng = self._engines.get(key)
        if eng is None:
            eng = self._new_engine(src, opts, cm, coupling, fields,
                                   quantities)
        else:
            self._engines.move_to_end(key)
        for k, v in self._options.items():
            if eng.options.get(k) != v:
                _native.check(_native.lib().mkh_sim_set_option(
                    eng.h, k.encode(), v), 'set_option')
                eng.options[k] = v
        old = self._engine
        if old is not None and self._dev_newer:
            # the latest state moves across on the device
            _native.check(_native.lib().mkh_sim_copy_state(eng.h, old.h),
                          'mkh_sim_copy_state')
            self._host_dirty = False
        else:
            self._host_dirty = True
        self._engines[key] = eng
        while len(self._engines) > 3:
            k0 = next(iter(self._engines))
            if self._engines[k0] is eng:
                break
            del self._engines[k0]
        self._engine = eng
        self._engine_key = key
        return eng, cm

    def _new_engine(self, src, opts, cm, coupling, fields, quantities):
        """Builds a device engine for one kernel text."""
        if _native.device_count() < 1:
            raise _native.DeviceError(
                'No HIP device found: SimulationOpenCL needs an AMD GPU.')
        tune = self._tuning
        prog = _program(src, opts, self._device)
        cfg = _native.SimConfig()
        cfg.precision = self._precision
        cfg.n_state = self._nstate
        cfg.n_field = len(fields)
        cfg.n_quant = len(quantities)
        cfg.i_vm = self._vm.index() if self._vm is not None else 0
        cfg.coupling = coupling
        cfg.driver = -1
        cfg.block_threads = self._threads()
        cfg.nx = self._nx
        cfg.ny = self._ny
        cfg.device = self._device
        cpt = tune.get('sheet_cpt') or _kernelgen.sheet_cpt(self._precision)
        cfg.sheet_cpt = cpt if _kernelgen.sheet_enabled(self._precision) \
            else 0
        cfg.chain_cpt = tune.get('chain_cpt') or \
            _kernelgen.chain_cpt(self._precision)
        cfg.adj_cpt = tune.get('adjoint_cpt') or \
            _kernelgen.adjoint_sheet_cpt(self._precision)
        eng = _Engine(prog, cfg)
        eng.options = {}
        eng.config_version = None
        eng.graph_version = None
        return eng

    def _configure(self, eng):
        """Sends configuration and state that changed since the engine last
        received them."""
        L = _native.lib()
        h = eng.h
        if eng.config_version != self._config_version:
            c = self._coupling()
            if c == _native.COUPLING_STENCIL:
                _native.check(L.mkh_sim_set_conductance(
                    h, self._gx, self._gy), 'set_conductance')
            elif c == _native.COUPLING_HETERO:
                gx = np.ascontiguousarray(self._gx_field, dtype=np.float64)
                gy = np.ascontiguousarray(
                    self._gy_field if self._gy_field is not None else [],
                    dtype=np.float64)
                _native.check(L.mkh_sim_set_conductance_field(
                    h, _native.ptr(gx, ctypes.c_double), len(gx),
                    _native.ptr(gy, ctypes.c_double), len(gy)),
                    'set_conductance_field')
            elif c == _native.COUPLING_GRAPH and \
                    eng.graph_version != self._graph_version:
                a = np.array(self._connections,
                             dtype=np.float64).reshape(-1, 3)
                ci = np.ascontiguousarray(a[:, 0], dtype=np.int64)
                cj = np.ascontiguousarray(a[:, 1], dtype=np.int64)
                _, iperm = self._cell_order()
                if iperm is not None:
                    ci = np.ascontiguousarray(iperm[ci])
                    cj = np.ascontiguousarray(iperm[cj])
                cg = np.ascontiguousarray(a[:, 2], dtype=np.float64)
                # identical copies held interleaved: the library keeps one
                # copy's rows when the list is the copies' lists in turn
                _native.check(L.mkh_sim_set_option(
                    h, b'graph_copies', int(self._copies)), 'set_option')
                _native.check(L.mkh_sim_set_connections(
                    h, _native.ptr(ci, ctypes.c_longlong),
                    _native.ptr(cj, ctypes.c_longlong),
                    _native.ptr(cg, ctypes.c_double), len(ci)),
                    'set_connections')
                eng.graph_version = self._graph_version
            perm, iperm = self._cell_order()
            paced = self._paced_cells
            if perm is not None and type(paced) == tuple:
                # a rectangle of a 1d graph is the cells x .. x + nx - 1;
                # in the device's order it is a list
                nx, ny, x, y = paced
                paced = list(range(max(0, x), min(self._ntotal, x + nx))) \
                    if y <= 0 < y + ny else []
            if not self._diffusion_enabled:
                _native.check(L.mkh_sim_set_pacing_all(h), 'pacing')
            elif type(paced) == tuple:
                nx, ny, x, y = paced
                _native.check(L.mkh_sim_set_pacing_rect(h, nx, ny, x, y),
                              'pacing')
            elif paced:
                cells = np.ascontiguousarray(paced, dtype=np.int64)
                if iperm is not None:
                    cells = np.ascontiguousarray(iperm[cells])
                _native.check(L.mkh_sim_set_pacing_list(
                    h, _native.ptr(cells, ctypes.c_longlong), len(cells)),
                    'pacing')
            else:
                _native.check(L.mkh_sim_set_pacing_none(h), 'pacing')
            if self._fields:
                fd = np.stack(list(self._fields.values()), axis=1)
                if perm is not None:
                    fd = fd[perm]
                fd = np.ascontiguousarray(fd.reshape(-1), dtype=np.float64)
                _native.check(L.mkh_sim_set_fields(
                    h, _native.ptr(fd, ctypes.c_double), len(fd)),
                    'set_fields')
            eng.config_version = self._config_version
        if self._host_dirty:
            st = np.asarray(self._state_arr, dtype=np.float64)
            perm, _ = self._cell_order()
            if perm is not None:
                st = st.reshape(self._ntotal, self._nstate)[perm].reshape(-1)
            st = np.ascontiguousarray(st)
            _native.check(L.mkh_sim_set_state(
                h, _native.ptr(st, ctypes.c_double), len(st)), 'set_state')
            self._host_dirty = False

    def _log_layout(self, log):
        """Returns (quantities, colmap, per-cell keys by column, global keys,
        logging_states)."""
        gtime = self._model.time()
        gpace = self._model.binding('pace')
        gnames = {}
        if gtime is not None:
            gnames[gtime.qname()] = 'time'
        if gpace is not None:
            gnames[gpace.qname()] = 'pace'
        quantities = []
        qindex = {}
        cols = []          # (key, q, cell)
        globals_ = []      # (key, 'time'|'pace')
        nx = self._nx
        for key in log.keys():
            kdims, kname = split_key(key)
            if not kdims:
                if kname in gnames:
                    globals_.append((key, gnames[kname]))
                    continue
                kdims = '0.'    # single cell logs without index (not used)
            var = self._model.get(kname)
            if var not in qindex:
                qindex[var] = len(quantities)
                quantities.append(var)
            idx = [int(i) for i in kdims[:-1].split('.')]
            cell = idx[0] + (idx[1] * nx if len(idx) > 1 else 0)
            cols.append((key, qindex[var], cell))
        colmap = np.full((len(quantities), self._ntotal), -1, dtype=np.int32)
        for col, (key, q, cell) in enumerate(cols):
            colmap[q, cell] = col
        logging_states = any(q.is_state() for q in quantities)
        return quantities, colmap, cols, globals_, logging_states

    def _run(self, duration, log, log_interval, report_nan, progress, msg):
        if duration < 0:
            raise ValueError('Simulation time can\'t be negative.')
        tmin = self._time
        tmax = tmin + duration
        g = []
        for label in self._global:
            v = self._model.binding(label)
            if v is not None:
                g.append(v.qname())
        log = prepare_log(
            log, self._model, dims=self._dims, global_vars=g,
            if_empty=LOG_STATE + LOG_BOUND,
            allowed_classes=LOG_STATE + LOG_INTER + LOG_BOUND,
            precision=self._precision)
        log_interval = 1e-9 if log_interval is None else float(log_interval)
        if log_interval <= 0:
            log_interval = 1e-9
        if progress is not None:
            from ._progress import ProgressReporter
            if not isinstance(progress, ProgressReporter):
                raise ValueError(
                    'The argument "progress" must be either a subclass of'
                    ' myokit.ProgressReporter or None.')

        arithmetic_error = False
        if duration > 0:
            quantities, colmap, cols, globals_, logging_states = \
                self._log_layout(log)
            values, times, paces, arithmetic_error = self._execute(
                tmin, tmax, quantities, colmap, len(cols), logging_states,
                log_interval, progress, msg)
            dtype = values.dtype
            self._fill_log(log, values, times, paces, cols, globals_, dtype)

        if report_nan and (arithmetic_error or log.has_nan()):
            self._report_nan(log)
        return log

    def _execute(self, tmin, tmax, quantities, colmap, ncols, logging_states,
                 log_interval, progress=None, msg=''):
        """Runs from tmin to tmax on the device. Returns (values[rows, ncols],
        times, paces, arithmetic_error)."""
        run = self._exec_begin(tmin, tmax, quantities, colmap, ncols,
                               logging_states, log_interval, progress)
        L = _native.lib()
        h = run.eng.h
        duration = tmax - tmin
        t = ctypes.c_double(tmin)
        status = ctypes.c_int(0)
        arithmetic_error = False
        ok = False
        try:
            if progress:
                with progress.job(msg):
                    r = 1.0 / duration if duration != 0 else 1
                    while status.value == 0:
                        code = L.mkh_sim_step(h, ctypes.byref(t),
                                              ctypes.byref(status))
                        if code == _native.MKH_E_NAN:
                            arithmetic_error = True
                            break
                        self._check_step(code)
                        if not progress.update(min((t.value - tmin) * r, 1)):
                            raise SimulationCancelledError()
            else:
                while status.value == 0:
                    code = L.mkh_sim_step(h, ctypes.byref(t),
                                          ctypes.byref(status))
                    if code == _native.MKH_E_NAN:
                        arithmetic_error = True
                        break
                    self._check_step(code)
            ok = True
        finally:
            values, times, paces = self._exec_end(run, ok)
        return values, times, paces, arithmetic_error

    def _exec_begin(self, tmin, tmax, quantities, colmap, ncols,
                    logging_states, log_interval, progress=None):
        """Prepares the engine for a run from tmin to tmax: logging, the
        step plan and the log buffers. Steps are then taken by mkh_sim_step
        (or, for two simulations in lockstep, mkh_ft_step) and the run is
        closed by :meth:`_exec_end`."""
        import types
        eng, cm = self._get_engine(quantities)
        if progress is not None and self._dev_newer:
            # Keep a host copy: a cancelled run leaves the state unchanged
            self._download_state()
        self._configure(eng)
        L = _native.lib()
        h = eng.h
        perm, iperm = self._cell_order()
        # the NaN test reads the caller's cell 0 (openclsim.c:1087-1089)
        nan_cell = 0 if iperm is None else int(iperm[0])
        if getattr(eng, 'nan_cell', 0) != nan_cell:
            _native.check(L.mkh_sim_set_nan_cell(h, nan_cell), 'nan_cell')
            eng.nan_cell = nan_cell
        if perm is not None:
            # columns of the device's cells; cached per column map
            key = id(colmap)
            hit = self._pcolmaps.get(key)
            if hit is None or hit[0] is not colmap:
                hit = (colmap, np.ascontiguousarray(
                    colmap.reshape(len(quantities), self._ntotal)[:, perm]))
                if len(self._pcolmaps) >= 8:
                    self._pcolmaps.clear()
                self._pcolmaps[key] = hit
            colmap = hit[1]
        _native.check(L.mkh_sim_set_logging(
            h, _native.ptr(np.ascontiguousarray(colmap.reshape(-1)),
                           ctypes.c_int),
            len(quantities), ncols, int(logging_states)), 'set_logging')
        events = np.ascontiguousarray(
            self._protocol.to_arrays() if self._protocol is not None
            else np.zeros((0, 5)), dtype=np.float64).reshape(-1)
        nrows = ctypes.c_longlong(0)
        code = L.mkh_sim_init(
            h, tmin, tmax, self._step_size,
            _native.ptr(events, ctypes.c_double), len(events) // 5,
            log_interval, ctypes.byref(nrows))
        if code == _native.MKH_E_PACING_SIMULTANEOUS:
            raise SimultaneousProtocolEventError(_native.last_error())
        _native.check(code, 'mkh_sim_init')
        nrows = nrows.value
        dtype = np.float64 if self._precision == DOUBLE_PRECISION \
            else np.float32
        shape = (nrows, max(ncols, 1))
        if nrows * shape[1] * np.dtype(dtype).itemsize >= (1 << 20):
            # page-locked: rows are copied while later steps compute
            values = _native.pinned.empty(shape, dtype)
        else:
            values = np.empty(shape, dtype=dtype)
        times = np.zeros(nrows, dtype=np.float64)
        paces = np.zeros(nrows, dtype=np.float64)
        _native.check(L.mkh_sim_set_log(
            h, values.ctypes.data_as(ctypes.c_void_p),
            _native.ptr(times, ctypes.c_double),
            _native.ptr(paces, ctypes.c_double)), 'mkh_sim_set_log')
        return types.SimpleNamespace(
            eng=eng, values=values, times=times, paces=paces,
            t0=_time.perf_counter())

    def _exec_end(self, run, ok):
        """Closes a run of :meth:`_exec_begin`; returns the logged rows
        (values, times, paces)."""
        L = _native.lib()
        h = run.eng.h
        L.mkh_sim_clean(h)
        self.last_run_stats = {
            'engine_seconds': _time.perf_counter() - run.t0,
            'cell_steps': None}
        if ok:
            self._dev_newer = True
        else:
            # Device state is partly advanced: next run uploads the host
            self._host_dirty = True
        rows = ctypes.c_longlong(0)
        L.mkh_sim_log_rows(h, ctypes.byref(rows))
        rows = rows.value
        return run.values[:rows], run.times[:rows], run.paces[:rows]

    def run_block(self, duration, variables=('membrane.V',),
                  log_interval=1.0, report_nan=True):
        """Runs like :meth:`run` with every cell of ``variables`` logged, but
        returns dense arrays instead of a DataLog: ``{'time': (rows,),
        'pace': (rows,), name: (rows, ny, nx) or (rows, nx)}``. This is the
        time x space view of ``myokit.DataBlock1d/2d``
        (``myokit/_datablock.py:87,787``) produced without per-cell keys."""
        if duration < 0:
            raise ValueError('Simulation time can\'t be negative.')
        log_interval = 1e-9 if log_interval is None else float(log_interval)
        if log_interval <= 0:
            log_interval = 1e-9
        quantities = []
        for v in variables:
            var = self._model.get(v) if not isinstance(v, Variable) else v
            if not (var.is_state() or var.is_intermediary() or (
                    var in self._bound_variables
                    and self._bound_variables[var] == 'idiff')):
                raise ValueError('Cannot log <' + var.qname() + '> per cell.')
            quantities.append(var)
        n = self._ntotal
        colmap = self._block_colmaps.get(len(quantities))
        if colmap is None:
            colmap = np.arange(len(quantities) * n, dtype=np.int32).reshape(
                len(quantities), n)
            colmap.setflags(write=False)
            self._block_colmaps[len(quantities)] = colmap
        tmin = self._time
        values, times, paces, err = self._execute(
            tmin, tmin + duration, quantities, colmap, colmap.size,
            any(q.is_state() for q in quantities), log_interval)
        self._time += duration
        shape = (self._ny, self._nx) if len(self._dims) == 2 else (self._nx,)
        out = {'time': times.astype(values.dtype), 'pace':
               paces.astype(values.dtype)}
        for k, var in enumerate(quantities):
            out[var.qname()] = values[:, k * n:(k + 1) * n].reshape(
                (len(times),) + shape)
        # As DataLog.has_nan: a NaN in the final logged row
        if report_nan and (err or (len(values) and np.isnan(values[-1]).any())):
            raise SimulationError(
                'Numerical error found in simulation logs.')
        return out

    def _check_step(self, code):
        if code == _native.MKH_E_PACING_SIMULTANEOUS:
            raise SimultaneousProtocolEventError(_native.last_error())
        _native.check(code, 'mkh_sim_step')

    def _fill_log(self, log, values, times, paces, cols, globals_, dtype):
        import collections
        new = {}
        for key, what in globals_:
            new[key] = (times if what == 'time' else paces).astype(dtype)
        for col, (key, q, cell) in enumerate(cols):
            new[key] = values[:, col]
        for key, arr in new.items():
            cur = log[key]
            if isinstance(cur, np.ndarray) and len(cur) == 0:
                collections.OrderedDict.__setitem__(log, key, arr)
            else:
                log.extend_column(key, arr)

    def _report_nan(self, log):
        """Raises SimulationError describing the first numerical error, as
        ``_run`` does (``openclsim.py:1151-1281``)."""
        fs = lambda x: float_str(x, precision=self._precision)  # noqa: E731
        txt = ['Numerical error found in simulation logs.']
        try:
            time, icell, var, value, states, bounds, d = self.find_nan(
                log, return_log=True)
            icell_str = '(' + ','.join(str(x) for x in icell) + ')'
            txt.append('Encountered numerical error at t=' + fs(time)
                       + ' in cell ' + icell_str + ' when ' + var + '='
                       + fs(value) + '.')
            is_paced = self.is_paced(*icell)
            is_paced_str = (', cell ' + icell_str + ' '
                            + ('IS' if is_paced else 'is NOT') + ' paced')
            vtime = self._model.time().qname()
            vpace = self._model.binding('pace')
            vpace = None if vpace is None else vpace.qname()
            vdiff = None
            if self._diffusion_enabled:
                vdiff = self._model.binding('diffusion_current')
                vdiff = None if vdiff is None else vdiff.qname()
            neighbors = self.neighbors(*icell)
            if len(self._dims) == 2:
                neighbors_str = ', '.join(
                    '(' + ','.join(str(i) for i in j) + ')' for j in neighbors)
            else:
                neighbors_str = ', '.join(str(j) for j in neighbors)

            def variables(k, when):
                txt.append('Simulation variables ' + when + ':')
                txt.append('  Time: ' + fs(bounds[k][vtime]))
                if vpace is not None:
                    txt.append('  Pacing variable: ' + fs(bounds[k][vpace])
                               + is_paced_str)
                if vdiff is not None:
                    txt.append('  Diffusion current: '
                               + fs(bounds[k][vdiff]))
                if neighbors:
                    txt.append('  Connected cells: ' + neighbors_str)

            txt.append('State during:')
            txt.append(self._model.format_state(states[0],
                                                precision=self._precision))
            variables(0, 'during')
            txt.append('Obtained ' + str(len(states))
                       + ' previous state(s).')
            if len(states) > 1:
                bound = {'time': bounds[1][vtime], 'pace': 0,
                         'diffusion_current': 0}
                if is_paced and vpace is not None:
                    bound['pace'] = bounds[1][vpace]
                if vdiff is not None:
                    bound['diffusion_current'] = bounds[1][vdiff]
                eval_error = None
                try:
                    derivs = self._model.evaluate_derivatives(
                        states[1], bound, self._precision,
                        ignore_errors=False)
                except NumericalError as ee:
                    derivs = self._model.evaluate_derivatives(
                        states[1], bound, self._precision,
                        ignore_errors=True)
                    eval_error = str(ee)
                txt.append('State before:')
                txt.append(self._model.format_state_derivatives(
                    states[1], derivs, self._precision))
                variables(1, 'before')
                dt = d.time()
                txt.append('Logged all variables for points: '
                           + ', '.join(fs(t) for t in dt))
                if len(dt) > 1:
                    txt.append('Non-finite valued variables at t='
                               + fs(dt[-2]))
                    for key, values in d.items():
                        if not np.isfinite(values[-2]):
                            txt.append('  ' + str(key) + ' = '
                                       + fs(values[-2]))
                if eval_error is not None:
                    txt.append('Error when evaluating derivatives:')
                    txt.append(eval_error)
        except FindNanError as e:
            txt.append('Unable to pinpoint source of NaN, an error occurred:')
            txt.append(str(e))
        raise SimulationError('\n'.join(txt))

    def find_nan(self, log, watch_var=None, safe_range=None,
                 return_log=False):
        """Searches for the origin of a bad value (NaN or inf) in a log of
        this simulation (``SimulationOpenCL.find_nan``,
        ``openclsim.py:395-702``).

        The log must hold the state of every cell and all bound variables.
        The first bad point (or, with ``watch_var`` and ``safe_range``, the
        first time that state leaves the range) is found by bisection, then
        the simulation is re-run from the logged point before it with a log
        interval ten times smaller, until the interval drops below 0.5 and
        the last re-run logs every step. Returns ``(time, icell, variable,
        value, states, bounds)`` (plus a log of every variable over the last
        points if ``return_log``): ``states[0]`` is the cell's state at the
        error, ``states[1..]`` up to three earlier ones, ``bounds`` the
        matching bound variables."""
        need = [self._model.binding(x).qname() for x in self._global
                if self._model.binding(x) is not None]
        need = prepare_log(LOG_STATE + LOG_BOUND, self._model,
                           dims=self._dims, global_vars=need)
        for key in need:
            if key not in log:
                raise FindNanError(
                    'Method requires a simulation log containing all states'
                    ' and bound variables. Missing variable <' + key + '>.')

        if watch_var is None:
            def bisect(ar, lo, hi):
                while True:
                    if not np.isfinite(ar[lo]):
                        return lo
                    md = lo + int(np.ceil(0.5 * (hi - lo)))
                    if md == hi:
                        return hi
                    if not np.isfinite(ar[md]):
                        hi = md
                    else:
                        lo = md

            def find_error_position(_log):
                ifirst = kfirst = None
                for key, ar in _log.items():
                    ar = np.asarray(ar)
                    if ifirst is None:
                        if not np.isfinite(ar[-1]):
                            kfirst = key
                            ifirst = bisect(ar, 0, len(ar) - 1)
                            if ifirst == 0:
                                break
                    elif not np.isfinite(ar[ifirst - 1]):
                        kfirst = key
                        ifirst = bisect(ar, 0, ifirst)
                        if ifirst == 0:
                            break
                return ifirst, kfirst
        else:
            try:
                watch_var = self._model.get(watch_var)
            except KeyError:
                raise FindNanError(
                    'Variable <' + str(watch_var) + '> not found.')
            if not watch_var.is_state():
                raise FindNanError('The watched variable must be a state.')
            try:
                lo, hi = safe_range
            except Exception:
                raise FindNanError(
                    'A safe range must be specified for the watched variable'
                    ' as a tuple (lower, upper).')
            if lo >= hi:
                raise FindNanError(
                    'The safe range must have a lower bound that is lower'
                    ' than the upper bound.')

            def find_error_position(_log):
                ifirst = kfirst = None
                post = '.' + watch_var.qname()
                for dims in dimco(*self._dims):
                    key = '.'.join(str(x) for x in dims) + post
                    ar = np.asarray(_log[key])
                    i = np.nonzero((ar < lo) | (ar > hi) | np.isnan(ar)
                                   | np.isinf(ar))[0]
                    if len(i):
                        i = i[0]
                        if ifirst is None or i < ifirst:
                            kfirst, ifirst = key, i
                        if i == 0:
                            break
                return ifirst, kfirst

        time_var = self._model.time().qname()
        states = list(self._model.states())

        def rerun(start, duration, flags, interval):
            state = []
            for dims in dimco(*self._dims):
                pre = '.'.join(str(x) for x in dims) + '.'
                for s in states:
                    state.append(float(log_[pre + s.qname()][start]))
            old_state, old_time = self._state, self._time
            self._state = state
            self._time = float(log_[time_var][start])
            try:
                return self.run(duration, log=flags, log_interval=interval,
                                report_nan=False)
            finally:
                self._state = old_state
                self._time = old_time

        log_ = log
        t = np.asarray(log_[time_var])
        dt = t[1] - t[0] if len(t) > 1 else 5
        while dt > 0:
            dt *= 0.1
            if dt < 0.5:
                dt = 0
            ifirst, kfirst = find_error_position(log_)
            if kfirst is None:
                raise FindNanError('Error condition not found in log.')
            if ifirst == 0:
                raise FindNanError(
                    'Unable to work with simulation logs where the error'
                    ' condition is met in the very first data point.')
            t = np.asarray(log_[time_var])
            log_ = rerun(ifirst - 1, t[ifirst] - t[ifirst - 1],
                         LOG_BOUND + LOG_STATE, dt)

        ifirst, kfirst = find_error_position(log_)
        if ifirst is None:
            ifirst = 0
            kfirst = next(iter(log_.keys()))
        ndims = len(self._dims)
        icell = [int(x) for x in kfirst.split('.')[0:ndims]]
        pre = '.'.join(str(x) for x in icell) + '.'

        def state_at(index):
            st = [float(log_[pre + v.qname()][index]) for v in states]
            b = {}
            for var in self._model.variables(bound=True):
                if var.binding() in self._global:
                    b[var.qname()] = float(log_[var.qname()][index])
                else:
                    b[var.qname()] = float(log_[pre + var.qname()][index])
            return st, b

        times, out_states, bounds = [], [], []
        for k in range(ifirst, ifirst - 4, -1):
            if k < 0:
                break
            times.append(float(log_[time_var][k]))
            st, b = state_at(k)
            out_states.append(st)
            bounds.append(b)
        var = self._model.get('.'.join(kfirst.split('.')[ndims:]))
        k = 1 if ifirst > 0 else 0
        value = out_states[k][var.index()] if var.is_state() else \
            bounds[k][var.qname()]
        var = var.qname()
        time = float(log_[time_var][ifirst])
        if not return_log:
            return time, icell, var, value, out_states, bounds
        d = rerun(max(0, ifirst - 3), len(times) * self._step_size,
                  LOG_BOUND + LOG_STATE + LOG_INTER, 0)
        return time, icell, var, value, out_states, bounds, d
