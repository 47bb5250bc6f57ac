#!/usr/bin/env python3
"""
Benchmark: cell-steps/s of the 2D 1024 x 1024 Luo-Rudy 1991 sheet in single
precision (Rush-Larsen, dt = 0.01 ms, gx = gy = 9.5, S1 planar stimulus on
the left 4 columns, V of every cell logged every 2 ms).

One bench "step" is one ``run_block`` call covering ``--segment-ms`` of
simulated time (default 20 ms = 2000 time steps of 1,048,576 cells). The
state stays resident on the device between steps; the log is copied back to
the host once per batch inside the timed region.

Prints one JSON line with the throughput, the HBM roofline of the dominant
kernel (algorithmic bytes per launch / mean launch time from HIP events),
and the CPU oracle timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CLOCK_GHZ = 2.4              # MI355X peak engine clock
SIMDS = 256 * 4              # 256 CUs x 4 SIMDs

# mk_sheet on this workload (FP32), from rocprofv3 counters of 200-step
# dispatches taken at commit 37cfa24
# (profiles/mk_sheet_lr1991_1024x1024_fp32.md), medians of 26 dispatches:
# - HBM traffic: FETCH_SIZE x2.0 + WRITE_SIZE x1.0 (factors of a 256 MiB
#   device copy, bench.py --copy-calibration), 153,900 x 2 + 446,500 KiB per
#   200 x 2^20 cell-steps;
# - SQ_INSTS_VALU 9.281e8 and SQ_INSTS_VALU_TRANS_F32 1.835e8 over 4,096
#   waves x 200 steps (wave-steps of 4 cells); a transcendental holds the
#   SIMD 8 cycles, any other VALU instruction 4 (SQ_ACTIVE_INST_VALU =
#   VALU - TRANS + 2 TRANS);
# - VALU busy (SQ_ACTIVE_INST_VALU x 4 / SIMDs / cycles per XCD): 0.87.
SHEET_PROFILE = {'file': 'profiles/mk_sheet_lr1991_1024x1024_fp32.md',
                 'commit': '37cfa24'}
SHEET_TRAFFIC_B_PER_CELL_STEP = ((153900 * 2 + 446500) * 1024.0
                                 / (200 * 2.0**20))
SHEET_VALU = 9.281e8 / (4096 * 200)
SHEET_TRANS = 1.835e8 / (4096 * 200)
SHEET_VALU_CYCLES_PER_WAVE_STEP = (SHEET_VALU - SHEET_TRANS) * 4 \
    + SHEET_TRANS * 8
SHEET_VALU_BUSY = 0.87
SHEET_CELLS_PER_LANE = 4


def copy_bandwidth(nbytes=2 << 30, reps=10):
    """Achievable device-to-device copy bandwidth in GB/s (bytes read plus
    bytes written per second), timed with HIP events around ``reps``
    hipMemcpy calls of ``nbytes``."""
    import ctypes
    hip = ctypes.CDLL('libamdhip64.so')
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    ms = ctypes.c_float()

    def ok(rc, what):
        if rc != 0:
            raise RuntimeError('%s failed (%d)' % (what, rc))
    ok(hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)), 'hipMalloc')
    try:
        ok(hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)),
           'hipMalloc')
        try:
            ok(hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)), 'hipMemset')
            ok(hip.hipEventCreate(ctypes.byref(e0)), 'hipEventCreate')
            ok(hip.hipEventCreate(ctypes.byref(e1)), 'hipEventCreate')
            ok(hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3), 'hipMemcpy')
            ok(hip.hipEventRecord(e0, None), 'hipEventRecord')
            for _ in range(reps):
                ok(hip.hipMemcpyAsync(b, a, ctypes.c_size_t(nbytes), 3, None),
                   'hipMemcpyAsync')
            ok(hip.hipEventRecord(e1, None), 'hipEventRecord')
            ok(hip.hipEventSynchronize(e1), 'hipEventSynchronize')
            ok(hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1),
               'hipEventElapsedTime')
            hip.hipEventDestroy(e0)
            hip.hipEventDestroy(e1)
        finally:
            hip.hipFree(b)
    finally:
        hip.hipFree(a)
    return 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9


S2_TIME = 320.0


def workload(nx, ny, precision):
    """The sheet with S1 (2 ms, left 4 columns, t = 0) and S2 (2 ms, lower
    left quadrant, t = 320 ms) in one protocol. The paced region is S1's
    until the run reaches 320 ms, where ``advance`` switches it to the
    quadrant, as a user of the reference does with run(320) followed by
    set_paced_cells (openclsim.py:1534-1629)."""
    import myokit_amd as myokit
    m = myokit.load_model('example')
    p = myokit.Protocol()
    p.schedule(1, 0, 2)          # S1 at t = 0, 2 ms
    p.schedule(1, S2_TIME, 2)    # S2 at t = 320 ms, 2 ms
    s = myokit.SimulationOpenCL(m, p, ncells=(nx, ny), rl=True,
                                precision=precision)
    s.set_step_size(0.01)
    s.set_conductance(9.5, 9.5)
    s.set_paced_cells(4, ny, 0, 0)
    return s


def advance(s, seg, log_interval=2.0):
    """One bench step: ``seg`` ms of the sheet from its current time. A
    step never crosses 320 ms (the segments divide it); at 320 ms the
    paced region becomes the lower-left quadrant for S2."""
    if abs(s.time() - S2_TIME) < 1e-9:
        s.set_paced_cells(s._nx // 2, s._ny // 2, 0, 0)
    return s.run_block(seg, ['membrane.V'], log_interval=log_interval)


def cpu_model():
    """The host CPU's model name (``/proc/cpuinfo``), or None."""
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(nx, ny, sample_ms, threads):
    """The CPU oracle (double-precision C, OpenMP) on the first
    ``sample_ms`` of the same workload."""
    import myokit_amd as myokit
    from oracle.cgen import OracleSimulation
    m = myokit.load_model('example')
    p = myokit.Protocol()
    p.schedule(1, 0, 2)
    o = OracleSimulation(m, p, ncells=(nx, ny), rl=True)
    o.dt = 0.01
    o.gx = o.gy = 9.5
    o.set_paced_cells(4, ny, 0, 0)
    o.threads = threads
    # warm up the build
    o.run(0.01, log_vars=['membrane.V'], log_interval=2.0, cells=[0])
    t0 = time.perf_counter()
    r = o.run(sample_ms, log_vars=['membrane.V'], log_interval=2.0)
    el = time.perf_counter() - t0
    return r['steps'] * nx * ny / el, r['steps'], el


# ---------------------------------------------------------------------------
# The other configurations of the scope table, each as its own workload
# (``--workload``): a bounded window of the configuration, timed like the
# sheet, with the algorithmic bytes per cell-step of its dominant kernel.
# ---------------------------------------------------------------------------

def _scales(rng, n):
    return np.exp(rng.uniform(np.log(0.5), np.log(2.0), n))


def _config1(precision=64):
    """200-cell LR1991 cable, FE, dt 0.005 ms, gx 9.5, cells 0-4 paced at
    1 Hz (2 ms, offset 0), V of every cell logged every 1 ms."""
    import myokit_amd as myokit
    m = myokit.load_model('example')
    s = myokit.SimulationOpenCL(m, myokit.pacing.blocktrain(1000, 2),
                                ncells=200, precision=precision)
    s.set_step_size(0.005)
    s.set_conductance(9.5)
    s.set_paced_cells(5)
    return s, 1.0


def _config3_cells(precision=32, n=65536, seed=3):
    """65,536 uncoupled LR1991 cells (diffusion off, all paced at 1 Hz),
    gNa, gCa (gsi) and gK scaled log-uniform in [0.5, 2] from a seed, FE,
    dt 0.005 ms, V logged every 0.5 ms."""
    import myokit_amd as myokit
    m = myokit.load_model('example')
    s = myokit.SimulationOpenCL(m, myokit.pacing.blocktrain(1000, 2),
                                ncells=n, diffusion=False,
                                precision=precision)
    s.set_step_size(0.005)
    rng = np.random.default_rng(seed)
    for name in ('ina.gNa', 'ica.gCa', 'ik.IK.gK'):
        s.set_field(name, m.get(name).rhs().eval() * _scales(rng, n))
    return s, 0.5


def _config3_cables(precision=32, ncables=256, seed=11):
    """256 LR1991 cables with lengths drawn uniformly in [50, 2000] from a
    seed, chains of g = 9.5 built with set_connections, the first 5 cells
    of each paced at 1 Hz, FE, dt 0.005 ms, V logged every 0.5 ms; all
    cables in one launch."""
    import myokit_amd as myokit
    m = myokit.load_model('example')
    rng = np.random.default_rng(seed)
    lengths = rng.integers(50, 2001, ncables)
    n = int(lengths.sum())
    rows, paced, off = [], [], 0
    for L in lengths:
        i = np.arange(off, off + L - 1)
        rows.append(np.stack([i, i + 1, np.full(L - 1, 9.5)], axis=1))
        paced += list(range(off, off + 5))
        off += L
    s = myokit.SimulationOpenCL(m, myokit.pacing.blocktrain(1000, 2),
                                ncells=n, precision=precision)
    s.set_step_size(0.005)
    s.set_connections(np.concatenate(rows))
    s.set_paced_cell_list(paced)
    return s, 0.5


def _smooth_field(rng, n, lo=0.7, hi=1.3):
    x = np.arange(n) / n
    f = np.zeros(n)
    for _ in range(8):
        f += rng.uniform(0.2, 1.0) * np.sin(
            2 * np.pi * rng.uniform(0.5, 6.0) * x + rng.uniform(0, 2 * np.pi))
    f = (f - f.min()) / (f.max() - f.min())
    return lo + (hi - lo) * f


def _config4(seed=7, gx=9.5):
    """1000-cell LR1991 cable, FP64, FE, dt 0.005 ms, cells 0-4 paced, gNa
    and gCa scale fields smooth in [0.7, 1.3] (8 seeded sinusoids)."""
    import myokit_amd as myokit
    m = myokit.load_model('example')
    rng = np.random.default_rng(seed)
    s = myokit.SimulationOpenCL(m, myokit.pacing.blocktrain(1000, 2),
                                ncells=1000, precision=64)
    s.set_step_size(0.005)
    s.set_conductance(gx)
    s.set_paced_cells(5)
    s.set_field('ina.gNa', 16.0 * _smooth_field(rng, 1000))
    s.set_field('ica.gCa', 0.09 * _smooth_field(rng, 1000))
    return s


def _config5(sets=32, n=262144, seed=21):
    """``sets`` parameter sets of a seeded random geometric graph of ``n``
    LR1991 cells in the unit square (mean degree 6, g uniform in [5, 12]),
    run as one batch (the disjoint union of one copy per set), FP32, RL,
    dt 0.01 ms, gNa, gCa and gK scaled log-uniform in [0.5, 2] per cell and
    set; cells within 0.05 of the square's left edge paced."""
    import myokit_amd as myokit
    from scipy.spatial import cKDTree
    m = myokit.load_model('example')
    rng = np.random.default_rng(seed)
    pts = rng.uniform(0, 1, (n, 2))
    r = np.sqrt(6.0 / (np.pi * n))
    pairs = cKDTree(pts).query_pairs(r, output_type='ndarray')
    g = rng.uniform(5, 12, len(pairs))
    base = np.stack([pairs[:, 0], pairs[:, 1], g], axis=1).astype(float)
    rows = np.concatenate([base + np.array([k * n, k * n, 0.0])
                           for k in range(sets)])
    left = np.nonzero(pts[:, 0] < 0.05)[0]
    s = myokit.SimulationOpenCL(m, myokit.pacing.blocktrain(1000, 2),
                                ncells=n * sets, precision=32, rl=True)
    s.set_step_size(0.01)
    s.set_connections(rows)
    s.set_paced_cell_list(np.concatenate([left + k * n
                                          for k in range(sets)]).tolist())
    for name in ('ina.gNa', 'ica.gCa', 'ik.IK.gK'):
        s.set_field(name, m.get(name).rhs().eval()
                    * _scales(rng, n * sets))
    return s, len(base)


def _time_blocks(s, seg, log_interval, steps, warmup):
    """Wall time of ``steps`` run_block calls of ``seg`` ms after
    ``warmup``, and the HIP-event kernel timing of one more."""
    for _ in range(warmup):
        s.run_block(seg, ['membrane.V'], log_interval=log_interval)
    t0 = time.perf_counter()
    for _ in range(steps):
        out = s.run_block(seg, ['membrane.V'], log_interval=log_interval)
    el = time.perf_counter() - t0
    assert np.all(np.isfinite(out['membrane.V']))
    s.set_engine_option('time_kernels', 1)
    s.engine_timing(reset=True)
    s.run_block(seg, ['membrane.V'], log_interval=log_interval)
    tm = s.engine_timing(reset=True)
    s.set_engine_option('time_kernels', 0)
    return el, tm


def _cpu_sample(o, sample_ms, threads):
    o.threads = threads
    o.run(o.dt, log_vars=['membrane.V'], log_interval=1e9, cells=[0])
    t0 = time.perf_counter()
    r = o.run(sample_ms, log_vars=['membrane.V'], log_interval=1e9,
              cells=[0])
    el = time.perf_counter() - t0
    return r['steps'] * o.n / el, r['steps'], el


def other_workload(args):
    """One JSON line for a configuration other than the 1024 x 1024
    sheet."""
    import myokit_amd as myokit
    from oracle.cgen import OracleSimulation
    w = args.workload
    threads = min(16, len(os.sched_getaffinity(0)))
    line = {'metric': 'cell-steps/s', 'unit': 'cell-steps/s', 'n_gpus': 1,
            'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
            'data': 'synthetic', 'steps': args.steps, 'warmup': args.warmup}
    oracle = None
    if w in ('cable', 'cells', 'cables'):
        prec = 64 if w == 'cable' else args.precision
        wb = 4 if prec == 32 else 8
        if w == 'cable':
            s, li = _config1(prec)
            seg = 100.0
            bpcs = 2 * 8 * wb
            desc = _config1.__doc__
            oracle = OracleSimulation(myokit.load_model('example'),
                                      myokit.pacing.blocktrain(1000, 2),
                                      ncells=200)
            oracle.gx = 9.5
            oracle.set_paced_cells(5, 1, 0, 0)
        elif w == 'cells':
            s, li = _config3_cells(prec)
            seg = 10.0
            bpcs = 2 * 8 * wb + 3 * wb
            desc = _config3_cells.__doc__
        else:
            s, li = _config3_cables(prec)
            seg = 10.0
            bpcs = 2 * 8 * wb + 2 * (4 + wb) + 2 * wb
            desc = _config3_cables.__doc__
        el, tm = _time_blocks(s, seg, li, args.steps, args.warmup)
        n = s._ntotal
        nsteps = int(round(seg / s._step_size))
        value = args.steps * nsteps * n / el
        avg_ms = tm['kernel_ms'] / max(tm['launches'], 1)
        cpl = tm['cell_steps'] / max(tm['launches'], 1)
        achieved = bpcs * cpl / (avg_ms * 1e-3) / 1e9
        line.update({
            'metric': 'cell-steps/s (%s)' % w, 'value': value,
            'ms_per_step': el / args.steps * 1e3,
            'dtype': 'f32' if prec == 32 else 'f64',
            'config': {'workload': ' '.join(desc.split()), 'cells': n,
                       'time_steps_per_step': nsteps,
                       'driver': s.engine_driver()},
            'roofline': {'bound': 'hbm', 'achieved': achieved,
                         'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                         'frac': achieved / HBM_PEAK_GBS, 'traffic': None,
                         'kernel': 'mk_' + s.engine_driver(),
                         'avg_launch_ms': avg_ms,
                         'bytes_per_cell_step': bpcs,
                         'cells_per_launch': cpl}})
        if w == 'cells':
            oracle = OracleSimulation(myokit.load_model('example'),
                                      myokit.pacing.blocktrain(1000, 2),
                                      ncells=n, diffusion=False)
            for var, v in s._fields.items():
                oracle.fields[var.qname()] = np.asarray(v, dtype=float)
        elif w == 'cables':
            oracle = OracleSimulation(myokit.load_model('example'),
                                      myokit.pacing.blocktrain(1000, 2),
                                      ncells=n)
            oracle.connections = np.asarray(s._connections)
            oracle.set_paced_cell_list(list(s._paced_cells))
        if oracle is not None and not args.no_cpu:
            oracle.dt = 0.005
            # a few seconds of host time each: the whole 1000 ms of the
            # 200-cell cable, 20 ms of the uncoupled cells, 10 ms of the
            # cables
            sample = {'cable': 1000.0, 'cells': 20.0, 'cables': 10.0}[w]
            cv, cs, cel = _cpu_sample(oracle, sample, threads)
            line['cpu_baseline'] = {
                'value': cv, 'unit': 'cell-steps/s', 'cores': threads,
                'kind': 'port', 'cpu': cpu_model(),
                'sample': 'CPU oracle (double-precision C, OpenMP) on the '
                          'first %g ms (%d time steps) of the same workload, '
                          '%.1f s' % (sample, cs, cel)}
        return line
    # adjoint workloads
    from myokit_amd.adjoint import TissueAdjoint
    if w == 'cable-adjoint':
        s = _config4()
        probes = np.linspace(25, 975, 20).astype(int)
        window, K = args.window_ms or 50.0, 200
        target = TissueAdjoint(_config4(101, 11.0), probes).trace(window)
        desc = ('config 4: ' + ' '.join(_config4.__doc__.split())
                + ' Loss over 20 probes against a trace of a hidden seeded'
                ' parameter set (other fields, gx 11); checkpoints every'
                ' 200 steps; window %g ms' % window)
        prec = 64
    else:
        s, nedges = _config5(sets=args.sets)
        n1 = s._ntotal // args.sets
        probes = np.concatenate([np.arange(0, n1, n1 // 4) + k * n1
                                 for k in range(args.sets)])
        window, K = args.window_ms or 2.0, 200
        target = np.full((int(round(window / 0.01)), len(probes)), -40.0)
        desc = ('config 5: ' + ' '.join(_config5.__doc__.split())
                + ' %d sets (%d cells, %d connections per set); loss over 4'
                ' probes per set; FP32 state and VJP, FP64 loss and gradient'
                ' reductions; window %g ms' % (args.sets, n1, nedges, window))
        prec = 32
    n = s._ntotal
    adj = TissueAdjoint(s, probes, checkpoint_every=K)
    nsteps = adj.n_steps(window)
    # forward alone: consecutive windows with one log row at the end of
    # each (the state stays on the device; a reset would time a full state
    # upload from the host as well)
    s.run_block(window, ['membrane.V'], log_interval=window)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.run_block(window, ['membrane.V'], log_interval=window)
    fwd = (time.perf_counter() - t0) / args.steps
    adj.loss_and_gradient(window, target)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        adj.loss_and_gradient(window, target)
    both = (time.perf_counter() - t0) / args.steps
    line.update({
        'metric': 'fwd+adjoint vs fwd time ratio (%s)' % w,
        'value': both / fwd, 'unit': 'x', 'higher_is_better': False,
        'ms_per_step': both * 1e3,
        'dtype': 'f32' if prec == 32 else 'f64',
        'config': {'workload': desc, 'cells': n, 'time_steps': nsteps,
                   'driver': s.engine_driver()},
        'adjoint': {'fwd_ms': fwd * 1e3, 'fwd_adjoint_ms': both * 1e3,
                    'phases_ms': adj.timing(),
                    'reverse_kernel': adj.reverse_kernel()[0],
                    'fwd_cell_steps_per_s': nsteps * n / fwd},
    })
    return line


def adjoint_leg(s, window_ms, checkpoint_every, reps, vjp_precision=None):
    """Times loss_and_gradient (forward + discrete adjoint, gradients for gx
    and gy) over ``window_ms`` from the sheet's current state, 20 probe
    cells on the diagonal, constant target."""
    from myokit_amd.adjoint import TissueAdjoint
    nx, ny = s._nx, s._ny
    probes = [(int((k + 0.5) * nx / 20), int((k + 0.5) * ny / 20))
              for k in range(20)]
    n_steps = None
    out = []
    for K in checkpoint_every:
        adj = TissueAdjoint(s, probes, checkpoint_every=K or 1 << 30,
                            vjp_precision=vjp_precision)
        if n_steps is None:
            n_steps = adj.n_steps(window_ms)
        target = np.full((n_steps, len(probes)), -80.0)
        adj.loss_and_gradient(window_ms, target)      # build, allocate
        t0 = time.perf_counter()
        for _ in range(reps):
            loss, g, dgx, dgy = adj.loss_and_gradient(window_ms, target)
        el = (time.perf_counter() - t0) / reps
        assert np.isfinite(loss) and np.isfinite(dgx) and np.isfinite(dgy)
        out.append((K or n_steps, el, adj.timing()))
        adj.release()
    return n_steps, out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--segment-ms', type=float, default=20.0)
    ap.add_argument('--nx', type=int, default=1024)
    ap.add_argument('--ny', type=int, default=1024)
    ap.add_argument('--precision', type=int, default=32)
    ap.add_argument('--cpu-sample-ms', type=float, default=10.0,
                    help='simulated time of the CPU oracle sample (10 ms of '
                    'the 1024x1024 sheet: ~12 s on 16 host threads)')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--adjoint-ms', type=float, default=20.0,
                    help='window of the fwd+adjoint timing (0: skip)')
    ap.add_argument('--adjoint-reps', type=int, default=2)
    ap.add_argument('--workload', default='sheet',
                    choices=['sheet', 'cable', 'cells', 'cables',
                             'cable-adjoint', 'graph-adjoint'],
                    help='sheet: the 1024x1024 sheet (default, the headline '
                    'metric); cable: config 1; cells and cables: config 3; '
                    'cable-adjoint: config 4; graph-adjoint: config 5')
    ap.add_argument('--window-ms', type=float, default=0.0,
                    help='adjoint workloads: simulated window')
    ap.add_argument('--sets', type=int, default=32,
                    help='graph-adjoint: parameter sets in the batch')
    ap.add_argument('--copy-calibration', action='store_true',
                    help='only fill 256 MiB and copy it device to device 4 '
                    'times (known traffic per dispatch, which scales the '
                    'HBM counters of a profile) and exit')
    args = ap.parse_args()

    if args.copy_calibration:
        copy_bandwidth(nbytes=256 << 20, reps=3)
        print('copy done: 1 fill and 4 copies of 256 MiB', flush=True)
        return

    if args.gpus != 1:
        print('bench.py runs on one GPU only: the engine has no multi-GPU'
              ' path (a sheet would shard by rows, one band per GPU, with'
              ' one halo row of V exchanged per time step).',
              file=sys.stderr)
        sys.exit(2)

    if args.workload != 'sheet':
        line = other_workload(args)
        line['n_gpus'] = 1
        print(json.dumps(line), flush=True)
        return

    s = workload(args.nx, args.ny, args.precision)
    seg = args.segment_ms
    n = args.nx * args.ny
    steps_per_seg = int(round(seg / 0.01))

    if abs(S2_TIME / seg - round(S2_TIME / seg)) > 1e-9:
        raise ValueError('--segment-ms must divide 320 ms (the S2 time).')
    for _ in range(args.warmup):
        advance(s, seg)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = advance(s, seg)
    el = time.perf_counter() - t0
    value = args.steps * steps_per_seg * n / el
    V = out['membrane.V']
    assert np.all(np.isfinite(V))

    # Roofline leg: one more segment with every launch bracketed by events
    s.set_engine_option('time_kernels', 1)
    s.engine_timing(reset=True)
    advance(s, seg)
    tm = s.engine_timing(reset=True)
    s.set_engine_option('time_kernels', 0)
    w = 4 if args.precision == 32 else 8
    bytes_per_cell_step = 2 * 8 * w       # state in + out, S = 8
    cells_per_launch = tm['cell_steps'] / max(tm['launches'], 1)
    avg_ms = tm['kernel_ms'] / max(tm['launches'], 1)
    achieved = bytes_per_cell_step * cells_per_launch / (avg_ms * 1e-3) / 1e9
    copy_gbs = copy_bandwidth()

    line = {
        'metric': 'cell-steps/s (2D %s LR1991 fp%d fwd); '
                  'fwd+adjoint vs fwd time ratio' % (
                      '1024²' if (args.nx, args.ny) == (1024, 1024)
                      else '%dx%d' % (args.nx, args.ny), args.precision),
        'value': value,
        'unit': 'cell-steps/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': el / args.steps * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32' if args.precision == 32 else 'f64',
        'data': 'synthetic',
        'config': {
            'workload': '2D sheet %dx%d LR1991 (example.mmt), Rush-Larsen, '
                        'dt=0.01 ms, gx=gy=9.5, S1 2 ms on the left 4 '
                        'columns at t=0, S2 2 ms on the lower-left quadrant '
                        'at t=320 ms, V of all cells logged every 2 ms; one '
                        'step = %g ms (%d time steps); timed steps cover '
                        '%g-%g ms' % (args.nx, args.ny, seg, steps_per_seg,
                                      args.warmup * seg,
                                      (args.warmup + args.steps) * seg),
            'cells': n, 'time_steps_per_step': steps_per_seg,
            'precision': args.precision,
            'driver': s.engine_driver(),
        },
        'roofline': {
            'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS,
            'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
            'traffic': None,
            'kernel': 'mk_' + s.engine_driver(), 'avg_launch_ms': avg_ms,
            'bytes_per_cell_step': bytes_per_cell_step,
            'cells_per_launch': cells_per_launch,
            'copy_gbs': copy_gbs,
            'frac_of_copy': achieved / copy_gbs,
        },
    }
    line['roofline']['achieved_is'] = (
        'algorithmic-equivalent bandwidth: %d B per cell-step (the state '
        'read and written once per time step) over the launch time; the '
        'persistent drivers keep the state in registers for a whole launch, '
        'so the HBM traffic they cause is roofline.traffic' % (
            bytes_per_cell_step))
    if (args.precision == 32 and s.engine_driver() == 'sheet'
            and (args.nx, args.ny) == (1024, 1024)):
        # Counter-derived figures of this kernel on this workload: HBM bytes
        # per launch and the VALU issue bound beside the HBM one
        steps_per_launch = cells_per_launch / n
        waves = n / (64 * SHEET_CELLS_PER_LANE)
        model_us = (waves / SIMDS * SHEET_VALU_CYCLES_PER_WAVE_STEP
                    / (CLOCK_GHZ * 1e3))
        meas_us = avg_ms * 1e3 / steps_per_launch
        traffic = SHEET_TRAFFIC_B_PER_CELL_STEP * cells_per_launch
        line['roofline']['traffic'] = traffic
        line['roofline']['traffic_gbs'] = traffic / (avg_ms * 1e-3) / 1e9
        line['roofline']['valu'] = {
            'bound_us_per_time_step': model_us,
            'measured_us_per_time_step': meas_us,
            'frac': model_us / meas_us,
            'busy_counter': SHEET_VALU_BUSY,
            'model': '(%.0f - %.0f) VALU x 4 + %.0f transcendental x 8 '
                     'cycles per wave-step of 4 cells, 4 SIMDs x 256 CUs, '
                     '2.4 GHz' % (SHEET_VALU, SHEET_TRANS, SHEET_TRANS),
        }
        line['roofline']['profile'] = SHEET_PROFILE
    if args.adjoint_ms > 0:
        fwd_ms = el / args.steps * 1e3 * args.adjoint_ms / seg
        n_adj, res = adjoint_leg(s, args.adjoint_ms, [0, 200],
                                 args.adjoint_reps)
        K0, t0, ph0 = res[0]
        line['adjoint'] = {
            'ratio': t0 * 1e3 / fwd_ms,
            'window_ms': args.adjoint_ms, 'time_steps': n_adj,
            'fwd_ms': fwd_ms, 'fwd_adjoint_ms': t0 * 1e3,
            'checkpoint_every': K0, 'phases_ms': ph0,
            'vjp_arithmetic': ('f32, removable singularities in f64'
                               if args.precision == 32 else 'f64'),
            'gradients': 'dL/dgx, dL/dgy; loss over 20 probe cells',
            'with_checkpoints_every_200': {
                'ratio': res[1][1] * 1e3 / fwd_ms,
                'fwd_adjoint_ms': res[1][1] * 1e3, 'phases_ms': res[1][2]},
        }
        if args.precision == 32:
            _, r64 = adjoint_leg(s, args.adjoint_ms, [0], args.adjoint_reps,
                                 vjp_precision=64)
            line['adjoint']['with_f64_vjp'] = {
                'ratio': r64[0][1] * 1e3 / fwd_ms,
                'fwd_adjoint_ms': r64[0][1] * 1e3, 'phases_ms': r64[0][2]}
    if not args.no_cpu:
        threads = min(16, len(os.sched_getaffinity(0)))
        cv, csteps, cel = cpu_baseline(args.nx, args.ny, args.cpu_sample_ms,
                                       threads)
        line['cpu_baseline'] = {
            'value': cv, 'unit': 'cell-steps/s', 'cores': threads,
            'kind': 'port', 'cpu': cpu_model(),
            'sample': 'CPU oracle (double-precision C, OpenMP) on the first '
                      '%g ms (%d time steps) of the same %dx%d workload, '
                      '%.1f s; the reference SimulationOpenCL on a CPU '
                      'OpenCL platform is unavailable (no PoCL)'
                      % (args.cpu_sample_ms, csteps, args.nx, args.ny, cel),
        }
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
