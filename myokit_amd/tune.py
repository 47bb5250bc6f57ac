"""
Autotuning of the launch configuration of a ``SimulationOpenCL``.

The engine picks a driver and launch shape by fixed rules (see
``SimulationOpenCL.set_tuning``). :func:`autotune` instead measures: it
runs a short window of the simulation's own workload from its current
state under each candidate configuration (state and time are restored
after every trial), keeps the fastest, applies it, and caches the choice
per model, geometry, precision, coupling and device, so later simulations
of the same kind can :func:`apply_cached` without measuring.

Candidates: the drivers that can run the geometry (uncoupled cells; one
workgroup per domain; one step per launch; the persistent sheet; the
persistent chains) crossed with threads per workgroup (128, 256, 512, 1024)
for the first three, cells per thread (tile heights 8, 16, 32) and register
budgets (waves per SIMD) for the sheet, cells per thread and register
budgets for the chains, and in double precision register budgets for the
one-step-per-launch driver. A candidate that fails to build or launch is
skipped.
"""
import hashlib
import json
import os
import time

from . import _native

CACHE_ENV = 'MYOKIT_AMD_TUNE_CACHE'


def _cache_path():
    p = os.environ.get(CACHE_ENV)
    if p:
        return p
    return os.path.join(os.path.expanduser('~'), '.cache', 'myokit_amd',
                        'tune.json')


def _key(sim):
    """Cache key: model equations, geometry, precision, coupling, fields,
    RL and the device."""
    h = hashlib.sha1()
    h.update(sim._model.code().encode())
    h.update(repr((sim._dims, sim._precision, sim._coupling(),
                   sorted(v.qname() for v in sim._fields),
                   sorted(s.qname() for s in sim._rl_states),
                   sim._connections is not None
                   and len(sim._connections) > 0)).encode())
    try:
        h.update(_native.device_info(sim._device)['name'].encode())
    except Exception:       # pragma: no cover
        pass
    return h.hexdigest()


def candidates(sim):
    """Launch configurations to try for this simulation, as keyword
    arguments of ``set_tuning``."""
    out = []
    c = sim._coupling()
    n = sim._ntotal
    bts = [128, 256, 512, 1024]
    if c == _native.COUPLING_NONE:
        drivers = ['cells', 'grid']
    elif c == _native.COUPLING_GRAPH:
        drivers = ['block', 'grid', 'chain']
    else:
        drivers = ['block', 'grid']
        if c == _native.COUPLING_STENCIL and sim._ny > 1:
            drivers.append('sheet')
    for d in drivers:
        if d == 'sheet':
            waves = (2, 3) if sim._precision == 64 else (3, 4, 5)
            cpts = (2, 4) if sim._precision == 64 else (2, 4, 8)
            for cpt in cpts:
                for w in waves:
                    out.append({'driver': 'sheet', 'block_threads': 256,
                                'sheet_cpt': cpt, 'sheet_waves': w})
        elif d == 'chain':
            waves = (2, 3) if sim._precision == 64 else (4, 5, 6)
            for cpt in (1, 2):
                for w in waves:
                    out.append({'driver': 'chain', 'chain_cpt': cpt,
                                'chain_waves': w})
        elif d == 'block' and n <= 64:
            out.append({'driver': 'block', 'block_threads': 64})
        elif d == 'grid':
            gws = (0, 4, 5, 6) if sim._precision == 64 else (0,)
            for bt in bts:
                for gw in gws:
                    out.append({'driver': d, 'block_threads': bt,
                                'grid_waves': gw})
        else:
            for bt in bts:
                out.append({'driver': d, 'block_threads': bt})
    return out


def _trial(sim, cfg, duration, reps):
    """Seconds per run of ``duration`` under ``cfg``, or None if the
    configuration cannot run here."""
    from ._err import DeviceError
    state, t0 = sim.state(), sim.time()
    try:
        sim.set_tuning(**cfg)
        sim.run(duration, log=[], report_nan=False)      # build, warm up
        sim.set_state(state)
        sim.set_time(t0)
        best = None
        for _ in range(reps):
            w = time.perf_counter()
            sim.run(duration, log=[], report_nan=False)
            w = time.perf_counter() - w
            best = w if best is None else min(best, w)
            sim.set_state(state)
            sim.set_time(t0)
        return best
    except (DeviceError, ValueError):
        return None
    finally:
        sim.set_state(state)
        sim.set_time(t0)


def autotune(sim, duration=None, reps=2, cache=True, verbose=False):
    """Measures every candidate configuration on a window of ``duration``
    (default: 200 time steps) of the simulation's own workload, applies the
    fastest and returns ``{'config': ..., 'seconds': ..., 'trials': [...]}``.
    The simulation's state and time are left unchanged."""
    if duration is None:
        duration = 200 * sim._step_size
    saved = sim.tuning()
    trials = []
    for cfg in candidates(sim):
        s = _trial(sim, cfg, duration, reps)
        trials.append((cfg, s))
        if verbose:
            print('tune', cfg, 'failed' if s is None else '%.4g s' % s,
                  flush=True)
    ok = [(s, i) for i, (c, s) in enumerate(trials) if s is not None]
    if not ok:
        sim.set_tuning(**saved)
        raise RuntimeError('No launch configuration could run.')
    s, i = min(ok)
    best = trials[i][0]
    sim.set_tuning(**best)
    if cache:
        _store(_key(sim), best)
    return {'config': best, 'seconds': s, 'trials': trials}


def _store(key, cfg):
    path = _cache_path()
    data = {}
    if os.path.isfile(path):
        try:
            with open(path) as f:
                data = json.load(f)
        except (OSError, ValueError):
            data = {}
    data[key] = cfg
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + '.tmp'
    with open(tmp, 'w') as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def apply_cached(sim):
    """Applies a cached choice for this kind of simulation; returns it, or
    None if there is none."""
    path = _cache_path()
    if not os.path.isfile(path):
        return None
    try:
        with open(path) as f:
            cfg = json.load(f).get(_key(sim))
    except (OSError, ValueError):
        return None
    if cfg:
        sim.set_tuning(**cfg)
    return cfg
