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
# dispatches (profiles/mk_sheet_lr1991_1024x1024_fp32.md):
# - HBM traffic per cell-step: FETCH_SIZE x2.0 + WRITE_SIZE x1.0 (factors of
#   a 256 MiB device copy), 307.6 + 340.8 MiB per 2.10e8 cell-steps;
# - per wave-step of 4 cells: 840 VALU instructions at 4 cycles and 224
#   transcendental ones (quarter rate) at 16.
SHEET_TRAFFIC_B_PER_CELL_STEP = (307.6 + 340.8) * 2.0**20 / (200 * 2.0**20)
SHEET_VALU = 840
SHEET_TRANS = 224
SHEET_VALU_CYCLES_PER_WAVE_STEP = SHEET_VALU * 4 + SHEET_TRANS * 16
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


def workload(nx, ny, precision):
    import myokit_amd as myokit
    m = myokit.load_model('example')
    p = myokit.Protocol()
    p.schedule(1, 0, 2)          # S1 at t = 0, 2 ms
    p.schedule(1, 320, 2)        # S2 at t = 320 ms, 2 ms
    s = myokit.SimulationOpenCL(m, p, ncells=(nx, ny), rl=True,
                                precision=precision)
    s.set_step_size(0.01)
    s.set_conductance(9.5, 9.5)
    s.set_paced_cells(4, ny, 0, 0)
    return s


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
    ap.add_argument('--cpu-sample-ms', type=float, default=0.5)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--adjoint-ms', type=float, default=20.0,
                    help='window of the fwd+adjoint timing (0: skip)')
    ap.add_argument('--adjoint-reps', type=int, default=2)
    args = ap.parse_args()

    if args.gpus != 1:
        print('bench.py runs on one GPU only: this build has no multi-GPU'
              ' path (see DESIGN.md for how the sheet would shard).',
              file=sys.stderr)
        sys.exit(2)

    s = workload(args.nx, args.ny, args.precision)
    seg = args.segment_ms
    n = args.nx * args.ny
    steps_per_seg = int(round(seg / 0.01))

    for _ in range(args.warmup):
        s.run_block(seg, ['membrane.V'], log_interval=2.0)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = s.run_block(seg, ['membrane.V'], log_interval=2.0)
    el = time.perf_counter() - t0
    value = args.steps * steps_per_seg * n / el
    V = out['membrane.V']
    assert np.all(np.isfinite(V))

    # Roofline leg: one more segment with every launch bracketed by events
    s.set_engine_option('time_kernels', 1)
    s.engine_timing(reset=True)
    s.run_block(seg, ['membrane.V'], log_interval=2.0)
    tm = s.engine_timing(reset=True)
    s.set_engine_option('time_kernels', 0)
    w = 4 if args.precision == 32 else 8
    bytes_per_cell_step = 2 * 8 * w       # state in + out, S = 8
    cells_per_launch = tm['cell_steps'] / max(tm['launches'], 1)
    avg_ms = tm['kernel_ms'] / max(tm['launches'], 1)
    achieved = bytes_per_cell_step * cells_per_launch / (avg_ms * 1e-3) / 1e9
    copy_gbs = copy_bandwidth()

    line = {
        'metric': 'cell-steps/s (2D 1024² LR1991 fp32 fwd); '
                  'fwd+adjoint vs fwd time ratio',
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
                        'dt=0.01 ms, gx=gy=9.5, S1 2 ms on left 4 columns, '
                        'V of all cells logged every 2 ms; one step = %g ms '
                        '(%d time steps)' % (args.nx, args.ny, seg,
                                              steps_per_seg),
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
    if (args.precision == 32 and s.engine_driver() == 'sheet'
            and (args.nx, args.ny) == (1024, 1024)):
        # Counter-derived figures of this kernel on this workload: HBM bytes
        # per launch and the VALU issue bound beside the HBM one
        steps_per_launch = cells_per_launch / n
        waves = n / (64 * SHEET_CELLS_PER_LANE)
        model_us = (waves / SIMDS * SHEET_VALU_CYCLES_PER_WAVE_STEP
                    / (CLOCK_GHZ * 1e3))
        meas_us = avg_ms * 1e3 / steps_per_launch
        line['roofline']['traffic'] = (SHEET_TRAFFIC_B_PER_CELL_STEP
                                       * cells_per_launch)
        line['roofline']['valu'] = {
            'bound_us_per_time_step': model_us,
            'measured_us_per_time_step': meas_us,
            'frac': model_us / meas_us,
            'model': '%d VALU x 4 + %d transcendental x 16 cycles per '
                     'wave-step of 4 cells, 4 SIMDs x 256 CUs, 2.4 GHz '
                     '(rocprofv3 SQ counters)' % (SHEET_VALU, SHEET_TRANS),
        }
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
            'kind': 'port',
            'sample': 'CPU oracle (double-precision C, OpenMP) on the first '
                      '%g ms (%d time steps) of the same %dx%d workload, '
                      '%.1f s' % (args.cpu_sample_ms, csteps, args.nx,
                                  args.ny, cel),
        }
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
