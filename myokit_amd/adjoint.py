"""
Gradients of a tissue simulation by the discrete adjoint (the reference has
no gradients; this is an addition on the same simulation object).

``TissueAdjoint(sim, probes)`` takes a configured ``SimulationOpenCL``. The
parameters are the fields set on it with ``set_field`` and the conductances
``gx``/``gy`` (for ``set_connections`` graphs: a common scale ``s`` of all
conductances). For a run of ``duration`` from the simulation's current state
and time, with one target row per time step,

    L = sum_{s=1..N} sum_p (V(t_s)[probe p] - target[s-1, p])^2,

``loss_and_gradient`` returns ``L``, ``dL/dfield`` (arrays shaped like the
field values), ``dL/dgx`` and ``dL/dgy``. The adjoint runs on the device:
forward with a checkpoint every ``checkpoint_every`` steps, then per
segment a recomputation of the states and the reverse steps. Loss and
gradients accumulate in double precision whatever the simulation precision.

``TissueLoss`` wraps this as a ``torch.autograd.Function`` so that the loss
can sit inside a PyTorch optimisation loop.
"""
import ctypes

import numpy as np

from . import _native
from ._err import SimultaneousProtocolEventError


class TissueAdjoint:
    """Discrete-adjoint gradients for a configured SimulationOpenCL."""

    def __init__(self, sim, probes, checkpoint_every=200,
                 vjp_precision=None):
        self.sim = sim
        # The VJP computes in the simulation's precision by default. In
        # single precision the equations with removable singularities are
        # differentiated in double (the quotient rule cancels there) and
        # the direct terms of mu are added last, which gives the accuracy
        # of the all-double VJP (vjp_precision=64) at about half the cost.
        from ._datalog import DOUBLE_PRECISION, SINGLE_PRECISION
        if vjp_precision is None:
            vjp_precision = sim._precision
        if vjp_precision not in (SINGLE_PRECISION, DOUBLE_PRECISION):
            raise ValueError('vjp_precision must be 32 or 64.')
        if sim._precision == DOUBLE_PRECISION:
            vjp_precision = DOUBLE_PRECISION
        self.vjp_precision = vjp_precision
        p = np.asarray(probes)
        if p.ndim == 2:
            # (x, y) pairs in 2D
            p = p[:, 0] + p[:, 1] * sim._nx
        self.probes = np.ascontiguousarray(p.reshape(-1), dtype=np.int64)
        if np.any(self.probes < 0) or np.any(self.probes >= sim._ntotal):
            raise ValueError('Probe cell out of range.')
        if len(np.unique(self.probes)) != len(self.probes):
            raise ValueError('Probe cells must be distinct.')
        self.checkpoint_every = int(checkpoint_every)
        if self.checkpoint_every < 1:
            raise ValueError('checkpoint_every must be at least 1.')

    def _events(self):
        sim = self.sim
        ev = sim._protocol.to_arrays().reshape(-1) if sim._protocol else \
            np.zeros(0)
        ev = np.ascontiguousarray(ev if len(ev) else np.zeros(1))
        nev = len(sim._protocol.events()) if sim._protocol else 0
        return ev, nev

    def n_steps(self, duration):
        """Number of time steps of a run of ``duration`` from the current
        time (the length the target must have)."""
        sim = self.sim
        ev, nev = self._events()
        n = ctypes.c_longlong(0)
        _native.check(_native.lib().mkh_plan(
            sim._time, sim._time + duration, sim._step_size,
            _native.ptr(ev, ctypes.c_double), nev, 1e300, None, None, None,
            None, 1 << 62, ctypes.byref(n)), 'mkh_plan')
        return n.value

    def _engine(self):
        sim = self.sim
        eng, cm = sim._get_engine([], adjoint=True,
                                  adjoint_precision=self.vjp_precision)
        sim._configure(eng)
        self._eng = eng
        return eng

    def _check(self, code, what):
        if code == _native.MKH_E_PACING_SIMULTANEOUS:
            raise SimultaneousProtocolEventError(_native.last_error())
        _native.check(code, what)

    def loss_and_gradient(self, duration, target):
        """Returns ``(loss, field_gradients, dgx, dgy)``; ``target`` has
        shape ``(n_steps(duration), n_probes)``."""
        sim = self.sim
        target = np.ascontiguousarray(target, dtype=np.float64)
        if target.ndim != 2 or target.shape[1] != len(self.probes):
            raise ValueError('Target must have shape (steps, n_probes).')
        eng = self._engine()
        ev, nev = self._events()
        nf = len(sim._fields)
        # page-locked: the gradients come back by DMA, without page faults
        gf = _native.pinned.empty((max(nf, 1) * sim._ntotal,), np.float64)
        gg = np.zeros(2)
        loss = ctypes.c_double(0)
        probes = self._device_probes()
        _, iperm = sim._cell_order()
        if getattr(eng, 'grad_order', None) is not iperm:
            # the device gathers the field gradients into the caller's cell
            # order; uploaded once per cell order and engine
            self._check(_native.lib().mkh_adj_set_cell_order(
                eng.h, None if iperm is None
                else _native.ptr(iperm, ctypes.c_longlong),
                0 if iperm is None else len(iperm)),
                'mkh_adj_set_cell_order')
            eng.grad_order = iperm
        self._check(_native.lib().mkh_adj_run(
            eng.h, sim._time, sim._time + duration, sim._step_size,
            _native.ptr(ev, ctypes.c_double), nev,
            _native.ptr(probes, ctypes.c_longlong), len(probes),
            _native.ptr(target, ctypes.c_double), target.shape[0],
            self.checkpoint_every, ctypes.byref(loss),
            _native.ptr(gf, ctypes.c_double),
            _native.ptr(gg, ctypes.c_double)), 'mkh_adj_run')
        shape = (sim._ny, sim._nx) if len(sim._dims) == 2 else (sim._nx,)
        grads = {}
        for k, var in enumerate(sim._fields.keys()):
            # views of the page-locked block, which goes back to its pool
            # once every one of them is gone (no copy, no fresh pages)
            grads[var.qname()] = gf[
                k * sim._ntotal:(k + 1) * sim._ntotal].reshape(shape)
        return loss.value, grads, gg[0], gg[1]

    def _device_probes(self):
        """The probe cells in the device's cell order."""
        _, iperm = self.sim._cell_order()
        if iperm is None:
            return self.probes
        return np.ascontiguousarray(iperm[self.probes], dtype=np.int64)

    def timing(self):
        """Wall time in ms of the last ``loss_and_gradient``'s phases:
        ``{'forward', 'recompute', 'reverse'}``. With the 'step-overlap'
        reverse kernel the recomputation runs during the reverse phase:
        'reverse' is then the whole phase and 'recompute' the part of it
        the second stream spent recomputing."""
        ms = np.zeros(3)
        eng = getattr(self, '_eng', None)
        if eng is None:
            return {'forward': 0.0, 'recompute': 0.0, 'reverse': 0.0}
        _native.check(_native.lib().mkh_adj_timing(
            eng.h, _native.ptr(ms, ctypes.c_double)), 'mkh_adj_timing')
        return {'forward': ms[0], 'recompute': ms[1], 'reverse': ms[2]}

    def reverse_kernel(self):
        """The reverse kernel of the last ``loss_and_gradient`` and its
        number of launches: ``('step' | 'step-overlap' | 'block' | 'sheet',
        launches)``, or ``(None, 0)`` before the first call. 'step-overlap'
        is the one-step kernel with each earlier segment recomputed on a
        second stream while the later one's reverse steps run (small
        domains)."""
        eng = getattr(self, '_eng', None)
        if eng is None:
            return None, 0
        kind, n = ctypes.c_int(-1), ctypes.c_longlong(0)
        _native.check(_native.lib().mkh_adj_reverse(
            eng.h, ctypes.byref(kind), ctypes.byref(n)), 'mkh_adj_reverse')
        names = {0: 'step', 1: 'block', 2: 'sheet', 3: 'step-overlap'}
        return names.get(kind.value), n.value

    def release(self):
        """Frees the device trajectory and work buffers kept between
        calls."""
        eng = getattr(self, '_eng', None)
        if eng is not None:
            _native.check(_native.lib().mkh_sim_set_option(
                eng.h, b'adj_release', 1), 'set_option')

    def trace(self, duration):
        """V at the probes after every time step of a run of ``duration``
        from the current state: shape ``(n_steps, n_probes)``, the quantity
        the loss compares with the target. Leaves the simulation's state and
        time unchanged."""
        sim = self.sim
        eng = self._engine()
        ev, nev = self._events()
        n = self.n_steps(duration)
        out = np.zeros((n, len(self.probes)))
        probes = self._device_probes()
        self._check(_native.lib().mkh_adj_trace(
            eng.h, sim._time, sim._time + duration, sim._step_size,
            _native.ptr(ev, ctypes.c_double), nev,
            _native.ptr(probes, ctypes.c_longlong), len(probes),
            _native.ptr(out, ctypes.c_double), n), 'mkh_adj_trace')
        return out


try:
    import torch

    class TissueLoss(torch.autograd.Function):
        """``loss = TissueLoss.apply(adjoint, duration, target, gx, gy,
        *field_tensors)`` with gradients for gx, gy and every field tensor
        (shaped like the values given to ``set_field``), by the discrete
        adjoint. Field tensors replace the simulation's fields, in the order
        of ``adjoint.sim._fields``."""

        @staticmethod
        def forward(ctx, adj, duration, target, gx, gy, *fields):
            sim = adj.sim
            if sim._diffusion_enabled and sim._connections is None and \
                    sim._gx_field is None:
                sim.set_conductance(float(gx), float(gy))
            names = [v.qname() for v in sim._fields.keys()]
            if len(fields) != len(names):
                raise ValueError('One tensor per field is required.')
            for name, f in zip(names, fields):
                sim.set_field(name, f.detach().cpu().numpy().astype(float))
            t = target.detach().cpu().numpy() if hasattr(target, 'detach') \
                else np.asarray(target)
            loss, grads, dgx, dgy = adj.loss_and_gradient(duration, t)
            ctx.dgx = dgx
            ctx.dgy = dgy
            ctx.grads = [torch.as_tensor(grads[n], dtype=f.dtype,
                                         device=f.device)
                         for n, f in zip(names, fields)]
            dev = fields[0].device if fields else 'cpu'
            return torch.tensor(loss, dtype=torch.float64, device=dev)

        @staticmethod
        def backward(ctx, gout):
            g = float(gout)
            out = [None, None, None,
                   torch.tensor(g * ctx.dgx, dtype=torch.float64),
                   torch.tensor(g * ctx.dgy, dtype=torch.float64)]
            out.extend(g * x for x in ctx.grads)
            return tuple(out)

except ImportError:     # pragma: no cover
    TissueLoss = None
