"""
FiberTissueSimulation: a fiber of cells leading into a rectangle of tissue,
each with its own cell model (``myokit.FiberTissueSimulation``,
``_sim/fiber_tissue.py:17-1251``, host loop ``_sim/fiber_tissue.c:
1003-1160``).

The fiber (``nfx x nfy`` cells) and the tissue (``ntx x nty``) are two
device simulations of the same kind as :class:`SimulationOpenCL`, each with
the kernel generated for its model, stepped in lockstep on one stream by the
device library (``mkh_ft_step``). Every time step computes, from the
potentials at ``t``:

* the stencil of each grid (zero-flux borders, conductances ``g_fiber`` and
  ``g_tissue``);
* the junction current ``g_fiber_tissue * (Vf - Vt)`` between the last cell
  of fiber row ``k`` and tissue cell ``(0, cty + k)``, with
  ``cty = int(0.5 * (nty - nfy))`` (``fiber_tissue.py:216-219``), added to
  the fiber cell's diffusion current and subtracted from the tissue cell's
  (``diff_step_fiber_tissue``, ``openclsim.cl:601-628``);
* the forward-Euler step of every cell; fiber cells with ``x < nx_paced``
  receive the protocol's pacing level, tissue cells none.

Time steps, log points and pacing follow ``SimulationOpenCL`` (the same
``sim_step`` rule); logs hold keys ``x.y.variable`` for each part.
"""
import contextlib
import ctypes

import numpy as np

from . import _native
from ._datalog import (DOUBLE_PRECISION, LOG_BOUND, LOG_INTER, LOG_NONE,
                       LOG_STATE, SINGLE_PRECISION, dimco, float_str,
                       prepare_log)
from ._err import (FindNanError, NumericalError, SimulationCancelledError,
                   SimulationError)

_GLOBAL = ('time', 'pace')


class FiberTissueSimulation:
    """Runs a fiber of cells leading up to a rectangular piece of tissue.

    Arguments as ``myokit.FiberTissueSimulation`` (``fiber_tissue.py:
    120-125``): ``fiber_model``, ``tissue_model``, ``protocol``,
    ``ncells_fiber=(nx, ny)``, ``ncells_tissue=(nx, ny)``, ``nx_paced``
    (width of the stimulus at the start of the fiber, over its full height),
    ``g_fiber=(gx, gy)``, ``g_tissue=(gx, gy)``, ``g_fiber_tissue``, ``dt``,
    ``precision`` and ``native_maths``. Both models need a state labelled
    ``membrane_potential`` and a variable bound to ``diffusion_current``,
    with the same units in both.
    """

    def __init__(self, fiber_model, tissue_model, protocol=None,
                 ncells_fiber=(128, 2), ncells_tissue=(128, 128),
                 nx_paced=5, g_fiber=(9, 6), g_tissue=(9, 6),
                 g_fiber_tissue=9, dt=0.005, precision=SINGLE_PRECISION,
                 native_maths=False):
        from .simulation import SimulationOpenCL
        fiber_model.validate()
        tissue_model.validate()
        for model, part in ((fiber_model, 'fiber'), (tissue_model, 'tissue')):
            if model.has_interdependent_components():
                cycles = '\n'.join(
                    '  ' + ' > '.join(x.name() for x in c)
                    for c in model.component_cycles())
                raise ValueError(
                    'This simulation requires models without interdependent'
                    ' components. Please restructure the ' + part + ' model'
                    ' and re-run. Cycles:\n' + cycles)

        # Sizes
        self._ncellsf = self._pair(ncells_fiber,
                                   'The fiber size must be a tuple (nx, ny).')
        self._ncellst = self._pair(ncells_tissue,
                                   'The tissue size must be a tuple (nx, ny).')
        self._ncellsf = [int(x) for x in self._ncellsf]
        self._ncellst = [int(x) for x in self._ncellst]
        if self._ncellsf[0] < 1 or self._ncellsf[1] < 1:
            raise ValueError('The fiber size must be at least (1, 1).')
        if self._ncellst[0] < 1 or self._ncellst[1] < 1:
            raise ValueError('The tissue size must be at least (1, 1).')
        if self._ncellsf[1] > self._ncellst[1]:
            raise ValueError(
                'The fiber y-dimension cannot exceed that of the tissue.')

        # Stimulus width
        nx_paced = int(nx_paced)
        if nx_paced < 0:
            raise ValueError(
                'The width of the stimulus pulse must be non-negative.')
        self._nx_paced = min(nx_paced, self._ncellsf[0])

        # Conductances
        self._gf = [float(x) for x in self._pair(
            g_fiber, 'The fiber conductivity must be a tuple (gx, gy).')]
        self._gt = [float(x) for x in self._pair(
            g_tissue, 'The tissue conductivity must be a tuple (gx, gy).')]
        self._gft = float(g_fiber_tissue)

        # Junction: the fiber's last column meets tissue column 0, centred
        self._cfx = self._ncellsf[0] - 1
        self._ctx = 0
        self._cty = int(0.5 * (self._ncellst[1] - self._ncellsf[1]))

        dt = float(dt)
        if dt <= 0:
            raise ValueError('The step size must be greater than zero.')
        self._step_size = dt
        if precision not in (SINGLE_PRECISION, DOUBLE_PRECISION):
            raise ValueError('Only single and double precision are supported.')
        self._precision = precision
        self._native_math = bool(native_maths)
        self._time = 0

        # Membrane potentials, diffusion currents and their units
        vms = []
        for model, part in ((fiber_model, 'fiber'), (tissue_model, 'tissue')):
            vm = model.label('membrane_potential')
            if vm is None:
                raise ValueError(
                    'This simulation requires the membrane potential variable'
                    ' to be labelled as "membrane_potential" in the ' + part
                    + ' model.')
            if not vm.is_state():
                raise ValueError(
                    'The variable labelled as membrane potential in the '
                    + part + ' model must be a state variable.')
            if model.binding('diffusion_current') is None:
                raise ValueError(
                    'This simulation requires a variable in the ' + part
                    + ' model to be bound to "diffusion_current" to pass'
                    ' current from one cell to the next.')
            vms.append(vm)
        for what, get in (
                ('membrane potential', lambda m, v: v),
                ('diffusion current',
                 lambda m, v: m.binding('diffusion_current'))):
            uf = get(fiber_model, vms[0]).unit()
            ut = get(tissue_model, vms[1]).unit()
            for u, part in ((uf, 'fiber'), (ut, 'tissue')):
                if u is None:
                    raise ValueError(
                        'The ' + part + ' model must specify a unit for the '
                        + what + '.')
            if uf != ut:
                raise ValueError(
                    'The ' + what + ' must have the same unit in the fiber'
                    ' and the tissue model: ' + str(uf) + ' vs ' + str(ut)
                    + '.')

        # The two parts
        self._f = SimulationOpenCL(
            fiber_model, protocol, ncells=tuple(self._ncellsf),
            precision=precision, native_maths=native_maths)
        self._t = SimulationOpenCL(
            tissue_model, protocol, ncells=tuple(self._ncellst),
            precision=precision, native_maths=native_maths)
        self._f.set_conductance(*self._gf)
        self._t.set_conductance(*self._gt)
        self._f.set_paced_cells(self._nx_paced, self._ncellsf[1], 0, 0)
        self._t.set_paced_cells(0, 0, 0, 0)
        for s in (self._f, self._t):
            s.set_step_size(dt)
            s.set_tuning(driver='grid')
        self._modelf = self._f._model
        self._modelt = self._t._model

    @staticmethod
    def _pair(value, msg):
        try:
            if len(value) != 2:
                raise ValueError(msg)
        except TypeError:
            raise ValueError(msg)
        return list(value)

    # ------------------------------------------------------------------
    # Shapes, time, step size, protocol
    # ------------------------------------------------------------------
    def fiber_shape(self):
        """The fiber grid's shape ``(ny, nx)``."""
        return (self._ncellsf[1], self._ncellsf[0])

    def tissue_shape(self):
        """The tissue grid's shape ``(ny, nx)``."""
        return (self._ncellst[1], self._ncellst[0])

    def set_step_size(self, step_size=0.005):
        step_size = float(step_size)
        if step_size <= 0:
            raise ValueError('Step size must be greater than zero.')
        self._step_size = step_size
        self._f.set_step_size(step_size)
        self._t.set_step_size(step_size)

    def step_size(self, step_size=0.005):
        """Returns the solver step size (the argument is ignored, as in the
        reference)."""
        return self._step_size

    def set_protocol(self, protocol=None):
        self._f.set_protocol(protocol)
        self._t.set_protocol(protocol)

    def set_time(self, time=0):
        self._time = float(time)

    def time(self):
        return self._time

    # ------------------------------------------------------------------
    # States
    # ------------------------------------------------------------------
    def _get(self, sim, dims, x, y, default, name):
        full = sim.default_state() if default else sim.state()
        if x is None:
            return list(full)
        if y is None:
            raise ValueError(
                name + '() must be called with no arguments or with both an'
                ' x and y argument.')
        x, y = int(x), int(y)
        if x < 0 or x >= dims[0]:
            raise IndexError('Given x-index out of range.')
        if y < 0 or y >= dims[1]:
            raise IndexError('Given y-index out of range.')
        ns = sim._model.count_states()
        k = x + y * dims[0]
        return list(full[k * ns:(k + 1) * ns])

    def _set(self, sim, dims, state, x, y, default, part):
        ns = sim._model.count_states()
        ntotal = dims[0] * dims[1]
        state = list(state)
        n = len(state)
        if n == ns * ntotal:
            new = state
        elif n != ns:
            raise ValueError(
                'Given state must have the same size as a single ' + part
                + ' cell state or a full simulation state')
        elif x is None:
            new = state * ntotal
        elif y is None:
            raise ValueError(
                'Must provide no coordinates, or both x and y coordinates.')
        else:
            x, y = int(x), int(y)
            if x < 0 or x >= dims[0]:
                raise IndexError('Given x-index out of range.')
            if y < 0 or y >= dims[1]:
                raise IndexError('Given y-index out of range.')
            new = list(sim.default_state() if default else sim.state())
            off = (x + y * dims[0]) * ns
            new[off:off + ns] = state
        if default:
            sim.set_default_state(new)
        else:
            sim.set_state(new)

    def fiber_state(self, x=None, y=None):
        """The fiber state (all cells, x changing first) or one cell's."""
        return self._get(self._f, self._ncellsf, x, y, False, 'fiber_state')

    def tissue_state(self, x=None, y=None):
        """The tissue state (all cells, x changing first) or one cell's."""
        return self._get(self._t, self._ncellst, x, y, False, 'tissue_state')

    def default_fiber_state(self, x=None, y=None):
        return self._get(self._f, self._ncellsf, x, y, True,
                         'default_fiber_state')

    def default_tissue_state(self, x=None, y=None):
        return self._get(self._t, self._ncellst, x, y, True,
                         'default_tissue_state')

    def set_fiber_state(self, state, x=None, y=None):
        self._set(self._f, self._ncellsf, state, x, y, False, 'fiber')

    def set_tissue_state(self, state, x=None, y=None):
        self._set(self._t, self._ncellst, state, x, y, False, 'tissue')

    def set_default_fiber_state(self, state, x=None, y=None):
        self._set(self._f, self._ncellsf, state, x, y, True, 'fiber')

    def set_default_tissue_state(self, state, x=None, y=None):
        self._set(self._t, self._ncellst, state, x, y, True, 'tissue')

    # ------------------------------------------------------------------
    # Running
    # ------------------------------------------------------------------
    def pre(self, duration, report_nan=True, progress=None,
            msg='Pre-pacing FiberTissueSimulation'):
        """Unlogged run; the final state becomes the new default state. The
        simulation time is not changed."""
        self._run(duration, LOG_NONE, LOG_NONE, 1, report_nan, progress, msg)
        self._f.set_default_state(self._f.state())
        self._t.set_default_state(self._t.state())

    def reset(self):
        """Time back to 0, state back to the default state."""
        self._time = 0
        self._f.reset()
        self._t.reset()

    def run(self, duration, logf=None, logt=None, log_interval=1.0,
            report_nan=True, progress=None,
            msg='Running FiberTissueSimulation'):
        """Runs for ``duration`` and returns ``(fiber_log, tissue_log)``
        (:class:`DataLog`; ``logf`` and ``logt`` as the ``log`` argument of
        ``SimulationOpenCL.run``, default: states and bound variables)."""
        r = self._run(duration, logf, logt, log_interval, report_nan,
                      progress, msg)
        self._time += duration
        return r

    def _globals(self, model):
        return [model.binding(x).qname() for x in _GLOBAL
                if model.binding(x) is not None]

    def _run(self, duration, logf, logt, log_interval, report_nan, progress,
             msg):
        import time as _time
        if duration < 0:
            raise ValueError('Simulation duration can\'t be negative.')
        tmin = self._time
        tmax = tmin + duration
        kw = dict(if_empty=LOG_STATE + LOG_BOUND,
                  allowed_classes=LOG_STATE + LOG_BOUND + LOG_INTER,
                  precision=self._precision)
        logf = prepare_log(logf, self._modelf, dims=self._ncellsf,
                           global_vars=self._globals(self._modelf), **kw)
        logt = prepare_log(logt, self._modelt, dims=self._ncellst,
                           global_vars=self._globals(self._modelt), **kw)
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
            parts = []
            for sim, log in ((self._f, logf), (self._t, logt)):
                parts.append((sim, log) + sim._log_layout(log))
            runs = []
            try:
                for sim, log, q, colmap, cols, g, ls in parts:
                    runs.append(sim._exec_begin(
                        tmin, tmax, q, colmap, len(cols), ls, log_interval,
                        progress))
                L = _native.lib()
                ft = ctypes.c_void_p()
                _native.check(L.mkh_ft_create(
                    runs[0].eng.h, runs[1].eng.h, self._ctx, self._cty,
                    self._gft, ctypes.byref(ft)), 'mkh_ft_create')
                t = ctypes.c_double(tmin)
                status = ctypes.c_int(0)
                ok = False
                t0 = _time.perf_counter()
                try:
                    r = 1.0 / duration
                    with (progress.job(msg) if progress
                          else contextlib.nullcontext()):
                        while status.value == 0:
                            code = L.mkh_ft_step(ft, ctypes.byref(t),
                                                 ctypes.byref(status))
                            if code == _native.MKH_E_NAN:
                                arithmetic_error = True
                                break
                            self._f._check_step(code)
                            if progress and not progress.update(
                                    min((t.value - tmin) * r, 1)):
                                raise SimulationCancelledError()
                    ok = True
                finally:
                    L.mkh_ft_destroy(ft)
                    self.last_run_stats = {
                        'engine_seconds': _time.perf_counter() - t0}
            except BaseException:
                for (sim, *_), run in zip(parts, runs):
                    sim._exec_end(run, False)
                raise
            dtype = np.float64 if self._precision == DOUBLE_PRECISION \
                else np.float32
            for (sim, log, q, colmap, cols, g, ls), run in zip(parts, runs):
                values, times, paces = sim._exec_end(run, ok)
                sim._fill_log(log, values, times, paces, cols, g, dtype)

        if report_nan and (arithmetic_error or logf.has_nan()
                           or logt.has_nan()):
            self._report_nan(logf, logt)
        return logf, logt

    # ------------------------------------------------------------------
    # NaN search (fiber_tissue.py:534-719)
    # ------------------------------------------------------------------
    def find_nan(self, logf, logt):
        """Searches for the origin of a NaN (or inf) in a pair of logs of
        this simulation holding every state and bound variable. Returns
        ``(part, time, icell, variable, value, states, bound)``: ``part`` is
        ``'fiber'`` or ``'tissue'``, ``states[0]`` the cell's state at the
        error and ``states[1..]`` up to three earlier ones, ``bound`` the
        matching bound variables."""
        for model, dims, log, part, extra in (
                (self._modelf, self._ncellsf, logf, 'fiber', ''),
                (self._modelt, self._ncellst, logt, 'tissue', ' at least')):
            need = prepare_log(LOG_STATE + LOG_BOUND, model, dims=dims,
                               global_vars=self._globals(model))
            for key in need:
                if key not in log:
                    raise FindNanError(
                        'Method requires a simulation log from the ' + part
                        + ' model containing all states and bound variables.'
                        ' Missing' + extra + ' variable <' + key + '>.')

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

        def find_error_position(log):
            ifirst = kfirst = None
            for key, ar in log.items():
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

        time_var = self._modelf.time().qname()

        def cell_states(model, dims, log, index):
            out = []
            for d in dimco(*dims):
                pre = '.'.join(str(x) for x in d) + '.'
                for s in model.states():
                    out.append(float(log[pre + s.qname()][index]))
            return out

        def relog(lf, lt, interval):
            ifirstf, kfirstf = find_error_position(lf)
            ifirstt, kfirstt = find_error_position(lt)
            if kfirstf is None and kfirstt is None:
                raise FindNanError('Error condition not found in logs.')
            elif kfirstf is None:
                ifirst = ifirstt
            elif kfirstt is None:
                ifirst = ifirstf
            elif ifirstf == 0 or ifirstt == 0:
                raise FindNanError(
                    'Unable to work with simulation logs where the error'
                    ' condition is met in the very first data point.')
            else:
                ifirst = min(ifirstf, ifirstt)
            istart = ifirst - 1
            statef = cell_states(self._modelf, self._ncellsf, lf, istart)
            statet = cell_states(self._modelt, self._ncellst, lt, istart)
            t = float(lf[time_var][istart])
            old = (self._f.state(), self._t.state(), self._time)
            self._f.set_state(statef)
            self._t.set_state(statet)
            self._time = t
            try:
                flags = LOG_BOUND + LOG_STATE
                return self.run(float(lf[time_var][ifirst]) - t, logf=flags,
                                logt=flags, log_interval=interval,
                                report_nan=False)
            finally:
                self._f.set_state(old[0])
                self._t.set_state(old[1])
                self._time = old[2]

        tf = np.asarray(logf[time_var])
        dt = tf[1] - tf[0] if len(tf) > 1 else 5
        while dt > 0:
            dt *= 0.1
            if dt < 0.5:
                dt = 0
            logf, logt = relog(logf, logt, dt)

        ifirstf, kfirstf = find_error_position(logf)
        ifirstt, kfirstt = find_error_position(logt)
        if kfirstt is None or (kfirstf is not None and kfirstf < kfirstt):
            part, ifirst, kfirst = 'fiber', ifirstf, kfirstf
            if ifirst is None:
                ifirst, kfirst = 0, next(iter(logf.keys()))
            model, log = self._modelf, logf
        else:
            part, ifirst, kfirst = 'tissue', ifirstt, kfirstt
            model, log = self._modelt, logt
        icell = [int(x) for x in kfirst.split('.')[0:2]]
        pre = '.'.join(str(x) for x in icell) + '.'

        def state_at(index):
            s = [float(log[pre + v.qname()][index]) for v in model.states()]
            b = {}
            for var in model.variables(bound=True):
                if var.binding() in _GLOBAL:
                    b[var.qname()] = float(log[var.qname()][index])
                else:
                    b[var.qname()] = float(log[pre + var.qname()][index])
            return s, b

        states, bound = [], []
        for k in range(ifirst, ifirst - 4, -1):
            if k < 0:
                break
            s, b = state_at(k)
            states.append(s)
            bound.append(b)
        var = model.get('.'.join(kfirst.split('.')[2:]))
        k = 1 if ifirst > 0 else 0
        value = states[k][var.index()] if var.is_state() else \
            bound[k][var.qname()]
        time = float(logf[time_var][ifirst])
        return part, time, icell, var.qname(), value, states, bound

    def _report_nan(self, logf, logt):
        """Raises SimulationError describing the first numerical error
        (``fiber_tissue.py:940-1017``)."""
        fs = lambda x: float_str(x, precision=self._precision)  # noqa: E731
        txt = ['Numerical error found in simulation logs.']
        try:
            part, time, icell, var, value, states, bounds = self.find_nan(
                logf, logt)
            model = self._modelt if part == 'tissue' else self._modelf
            txt.append(
                'Encountered numerical error in ' + part + ' simulation at'
                ' t =' + fs(time) + ' in cell ('
                + ','.join(str(x) for x in icell) + ') when ' + var + '= '
                + fs(value) + '.')
            vtime = model.time().qname()
            vdiff = model.binding('diffusion_current').qname()
            vpace = model.binding('pace')
            vpace = None if vpace is None else vpace.qname()
            txt.append('State during:')
            txt.append(model.format_state(states[0],
                                          precision=self._precision))
            txt.append('Simulation variables during:')
            txt.append('  Time:' + fs(bounds[0][vtime]))
            if vpace is not None:
                txt.append('  Pacing variable:' + fs(bounds[0][vpace]))
            txt.append('  Diffusion current:' + fs(bounds[0][vdiff]))
            txt.append('Obtained ' + str(len(states)) + ' previous state(s).')
            if len(states) > 1:
                bound = {
                    'time': bounds[1][vtime],
                    'pace': 0 if vpace is None else bounds[1][vpace],
                    'diffusion_current': bounds[1][vdiff],
                }
                eval_error = None
                try:
                    derivs = model.evaluate_derivatives(
                        states[1], bound, self._precision,
                        ignore_errors=False)
                except NumericalError as ee:
                    derivs = model.evaluate_derivatives(
                        states[1], bound, self._precision,
                        ignore_errors=True)
                    eval_error = str(ee).strip()
                txt.append('State before:')
                txt.append(model.format_state_derivatives(
                    states[1], derivs, self._precision))
                txt.append('Simulation variables during:')
                txt.append('  Time:' + fs(bounds[1][vtime]))
                if vpace is not None:
                    txt.append('  Pacing variable:' + fs(bounds[0][vpace]))
                txt.append('  Diffusion current:' + fs(bounds[1][vdiff]))
                if eval_error is not None:
                    txt.append('Error when evaluating derivatives:')
                    txt.append(eval_error)
        except FindNanError as e:
            txt.append('Unable to pinpoint source of NaN, an error occurred:')
            txt.append(str(e))
        raise SimulationError('\n'.join(txt))
