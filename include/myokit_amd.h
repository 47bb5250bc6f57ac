/*
 * myokit_amd.h -- C ABI of the MI355X tissue-simulation library
 * (libmyokit_amd.so).
 *
 * This is the boundary that replaces the per-run CPython extension which
 * myokit's SimulationOpenCL compiles from myokit/_sim/openclsim.c:
 *
 *   reference (myokit/_sim/openclsim.c)           this library
 *   ---------------------------------------      -----------------------------
 *   clCreateProgramWithSource/clBuildProgram      mkh_program_build
 *     (:815-839)
 *   sim_init(21 Python args) (:309-1031)          mkh_sim_create + mkh_sim_set_*
 *                                                 + mkh_sim_init + mkh_sim_set_log
 *   sim_step() -> time / ArithmeticError          mkh_sim_step
 *     (:1036-1211)
 *   sim_clean() (:216-292, :1216-1221)            mkh_sim_clean
 *   final state read into state_out (:1185-1190)  mkh_sim_get_state
 *   mcl_select_device / mcl_info (mcl.h:265,823)  mkh_device_count/_info
 *
 * All entry points take plain pointers and sizes. They return 0 on success
 * or a negative MKH_E_* code; mkh_last_error() then holds a message. A
 * simulation handle owns its device memory. Handles are independent: several
 * may exist per process and per device. A single handle is not thread-safe.
 *
 * Host arrays use the reference's layouts: state is cell-major
 * (state[cell * n_state + i], cell = x + y * nx, openclsim.cl:316-319); field
 * data is cell-major and field-minor (openclsim.py:1082-1089). On the device,
 * state and fields are stored structure-of-arrays.
 */
#ifndef MYOKIT_AMD_H
#define MYOKIT_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes */
#define MKH_OK 0
#define MKH_E_ARG -1          /* invalid argument */
#define MKH_E_HIP -2          /* HIP runtime error (message has the name) */
#define MKH_E_COMPILE -3      /* hiprtc failed to compile the kernel */
#define MKH_E_STATE -4        /* call out of order (e.g. step before init) */
#define MKH_E_NAN -5          /* NaN in cell 0 at a log point: run halted */
#define MKH_E_MEMORY -6       /* allocation failed */
/* Pacing errors, same meaning as the ESys flags of myokit/_sim/pacing.h:47-64 */
#define MKH_E_PACING_NEGATIVE_TIME -40
#define MKH_E_PACING_SIMULTANEOUS -50
#define MKH_E_PACING_INVALID -20

/* Coupling modes (which diffusion variant the kernel computes) */
#define MKH_COUPLING_NONE 0      /* diffusion disabled, or a single cell */
#define MKH_COUPLING_STENCIL 1   /* homogeneous 1D/2D grid (diff_step, openclsim.cl:384) */
#define MKH_COUPLING_HETERO 2    /* conductance fields (diff_hetero, :452) */
#define MKH_COUPLING_GRAPH 3     /* set_connections (diff_arb_*, :499-573) */

/* Execution drivers */
#define MKH_DRIVER_CELLS 0  /* one thread per uncoupled cell, all steps per launch */
#define MKH_DRIVER_BLOCK 1  /* one workgroup per domain, all steps per launch */
#define MKH_DRIVER_GRID 2   /* whole grid, one fused step per launch */
#define MKH_DRIVER_SHEET 3  /* 2D sheet, persistent tiles, all steps per launch */
#define MKH_DRIVER_CHAIN 4  /* graphs of disjoint chains (cables) laid end to
                               end: persistent one-wave tiles, all steps per
                               launch */

/* Simulation geometry and kernel layout. */
typedef struct mkh_sim_config {
    int precision;         /* 32 or 64 (myokit.SINGLE/DOUBLE_PRECISION) */
    int n_state;           /* states per cell */
    int n_field;           /* per-cell scalar fields */
    int n_quant;           /* per-cell logged quantities compiled into the kernel */
    int i_vm;              /* index of the membrane potential state */
    int coupling;          /* MKH_COUPLING_* */
    int driver;            /* MKH_DRIVER_* */
    int block_threads;     /* threads per workgroup */
    long long nx, ny;      /* grid size (1D: ny = 1) */
    int device;            /* HIP device ordinal */
    int sheet_cpt;         /* cells per thread of the sheet driver (MK_SCPT in
                              the kernel source), 0 = no sheet driver */
    int chain_cpt;         /* cells per thread of the chain driver (MK_CCPT in
                              the kernel source), 0 = no chain driver */
    int adj_cpt;           /* cells per thread of the persistent reverse sheet
                              (MK_ACPT in the kernel source: tiles of
                              64 x 4 adj_cpt cells), 0 = no reverse sheet */
} mkh_sim_config;

/* Library */
const char* mkh_version(void);
const char* mkh_last_error(void);
int mkh_device_count(int* n);
int mkh_device_info(int device, char* name, int name_len,
                    long long* total_mem, int* n_cu);

/* Kernel program: compile generated HIP source with hiprtc for the device's
 * architecture (gfx950) and load it. Replaces clCreateProgramWithSource +
 * clBuildProgram (openclsim.c:815-839). If cache_dir is non-NULL, code
 * objects are cached there by a hash of source and options. build_log (may
 * be NULL) receives the compiler log. */
int mkh_program_build(const char* source, const char* options,
                      const char* cache_dir, int device, void** program,
                      char* build_log, int build_log_len);
int mkh_program_free(void* program);

/* Simulation object (device buffers for one tissue). */
int mkh_sim_create(void* program, const mkh_sim_config* cfg, void** sim);
int mkh_sim_destroy(void* sim);

/* State in and out (host AoS doubles, n = n_cells * n_state). Replaces the
 * PyList conversions of openclsim.c:512-524 and :1185-1190. */
int mkh_sim_set_state(void* sim, const double* state, long long n);
int mkh_sim_get_state(void* sim, double* state, long long n);

/* Copies the current device state of simulation src into dst, device to
 * device (both of the same precision and geometry, e.g. engines of one
 * model compiled for different log selections). No counterpart in the
 * reference, whose state lives in host lists between runs. */
int mkh_sim_copy_state(void* dst, void* src);

/* FiberTissueSimulation (_sim/fiber_tissue.py:17, its sim_step
 * _sim/fiber_tissue.c:1003-1160 and the junction kernel
 * diff_step_fiber_tissue, _sim/openclsim.cl:601-628). `fiber` and `tissue`
 * are two simulations of one precision with stencil coupling and the grid
 * driver; fiber row k's last cell connects to tissue cell (ctx, cty + k)
 * with conductance gft. Both are initialised with mkh_sim_init for the
 * same run (and given their log buffers); each mkh_ft_step then advances
 * both by one batch in lockstep, like mkh_sim_step (status 0 more, 1 done;
 * MKH_E_NAN when the NaN check of either part halts the run, both logs
 * then ending at the same row). */
int mkh_ft_create(void* fiber, void* tissue, long long ctx, long long cty,
                  double gft, void** ft);
int mkh_ft_step(void* ft, double* t, int* status);
int mkh_ft_destroy(void* ft);

/* Per-cell fields, cell-major / field-minor (openclsim.py:1082-1089). */
int mkh_sim_set_fields(void* sim, const double* data, long long n);

/* Conductances (openclsim.py:1286-1449). Connections are (i, j, g) with
 * i < j, validated by the caller; summation order per cell follows the
 * order of the list. */
int mkh_sim_set_conductance(void* sim, double gx, double gy);
int mkh_sim_set_conductance_field(void* sim, const double* gx, long long ngx,
                                  const double* gy, long long ngy);
int mkh_sim_set_connections(void* sim, const long long* i,
                            const long long* j, const double* g,
                            long long n);

/* Independent domains for MKH_DRIVER_BLOCK: domain d owns cells
 * [offsets[d], offsets[d+1]); n_domains + 1 offsets. Each domain is a 1D
 * chain (stencil/hetero) or a connected set (graph). Without this call the
 * whole grid is one domain. */
int mkh_sim_set_domains(void* sim, const long long* offsets,
                        long long n_domains);

/* Paced cells (calculate_pacing, openclsim.cl:249-280). */
int mkh_sim_set_pacing_all(void* sim);
int mkh_sim_set_pacing_none(void* sim);
int mkh_sim_set_pacing_rect(void* sim, long long nx, long long ny,
                            long long x, long long y);
int mkh_sim_set_pacing_list(void* sim, const long long* cells, long long n);

/* Logging: colmap[q * n_cells + cell] is the log column that per-cell
 * quantity q of that cell is written to, or -1. n_cols is the number of
 * per-cell columns. nan_check enables the reference's NaN test on state 0
 * of cell 0 (see mkh_sim_set_nan_cell) at every log point
 * (openclsim.c:1079-1090). */
int mkh_sim_set_logging(void* sim, const int* colmap, long long n_quant,
                        int n_cols, int nan_check);

/* The cell whose state 0 the NaN test reads (default 0). The reference tests
 * the caller's cell 0 (rvec_state[0], openclsim.c:1087-1089); a caller that
 * holds its cells on the device in another order (e.g. a renumbered
 * connection graph) passes the device index of its cell 0. */
int mkh_sim_set_nan_cell(void* sim, long long cell);

/* Begin a run from tmin to tmax with step dt (sim_init, openclsim.c:309).
 * events: n_events rows of (level, start, duration, period, multiplier),
 * in protocol order. Returns the number of log rows the run will produce
 * in *n_log_rows (the time-step rule is deterministic). */
int mkh_sim_init(void* sim, double tmin, double tmax, double dt,
                 const double* events, long long n_events,
                 double log_interval, long long* n_log_rows);

/* Host destinations for the log: values[row * n_cols + col] (precision of
 * the simulation), times[row] and paces[row] as doubles holding the
 * precision-cast engine time and pacing level. Any may be NULL if unused. */
int mkh_sim_set_log(void* sim, void* values, double* times, double* paces);

/* Advance by one bounded batch of steps (sim_step, openclsim.c:1036).
 * *time receives the engine time; *status is 0 (more to do), 1 (finished)
 * or 2 (halted on NaN, in which case MKH_E_NAN is also returned). */
int mkh_sim_step(void* sim, double* time, int* status);

/* Release the per-run state (sim_clean, openclsim.c:216). Device state
 * stays resident. */
int mkh_sim_clean(void* sim);

/* Tuning and measurement options:
 *   "batch_steps"   max steps per mkh_sim_step call
 *   "log_rows"      device log buffer rows per batch
 *   "time_kernels"  1: bracket every launch with HIP events
 *   "block_threads" threads per workgroup for the next runs
 *   "driver"        force a driver (MKH_DRIVER_*), -1 = automatic
 *   "adj_release"   1: free the adjoint's trajectory and work buffers
 *   "adj_sheet"     0: on a persistent sheet, take the adjoint's reverse
 *                   sweep one launch per step instead of one persistent
 *                   launch per segment (default 1)
 *   "adj_overlap"   0: recompute earlier segments in line with the reverse
 *                   steps, also where a second stream could overlap them
 *                   (default 1)
 *   "graph_copies"  K: the next graph is K identical copies held
 *                   interleaved (cell p of copy k is cell p * K + k), the
 *                   copies' connection lists one after the other; if every
 *                   copy's list is the first's shifted, with the same
 *                   conductances, one copy's rows serve all (default 1) */
int mkh_sim_set_option(void* sim, const char* key, long long value);

/* Kernel timing since the last reset (time_kernels = 1): total milliseconds
 * inside kernels, number of launches and cell-steps computed. */
int mkh_sim_get_timing(void* sim, double* kernel_ms, long long* launches,
                       long long* cell_steps);
int mkh_sim_reset_timing(void* sim);

/* Discrete adjoint of a run. No counterpart in the reference (myokit has no
 * gradients); for callers that fit per-cell fields and conductances to
 * recorded potentials. Runs tmin..tmax from the simulation's current state
 * (left unchanged) with the loss
 *     L = sum_{s=1..N} sum_p (V(t_s)[probes[p]] - target[(s-1) * n_probes + p])^2
 * over the N time steps of the run (one target row per step), and returns L,
 * dL/dfield as grad_fields[k * n_cells + c], and dL/dgx, dL/dgy (graphs:
 * dL/ds for all conductances scaled by s) in grad_g[0..1]. The kernel must
 * be generated with the adjoint. The forward pass keeps a checkpoint every
 * checkpoint_every steps and all states of the last segment; earlier
 * segments are recomputed from their checkpoints during the reverse pass
 * (device memory: (checkpoint_every + 1) states plus one per checkpoint).
 * Loss and gradients accumulate in double precision. */
int mkh_adj_run(void* sim, double tmin, double tmax, double dt,
                const double* events, long long n_events,
                const long long* probes, long long n_probes,
                const double* target, long long n_target_rows,
                int checkpoint_every, double* loss, double* grad_fields,
                double* grad_g);

/* Page-locked host memory for log destinations (mkh_sim_set_log): rows
 * copied into it leave the device while later steps compute. The reference
 * appends each logged value to Python lists instead (openclsim.c:1120-1131);
 * these two calls have no counterpart there. */
int mkh_host_alloc(long long bytes, void** ptr);
int mkh_host_free(void* ptr);

/* Wall time in ms of the last mkh_adj_run's phases: ms3[0] forward with
 * checkpoints, ms3[1] recomputation of earlier segments, ms3[2] reverse
 * steps (with MKH_ADJ_REV_STEP_OVERLAP the recomputation runs during the
 * reverse steps: ms3[2] is then the whole reverse phase, ms3[1] the part
 * of it the second stream spent recomputing). No counterpart in the
 * reference. */
int mkh_adj_timing(void* sim, double* ms3);

/* The reverse kernel the last mkh_adj_run used: MKH_ADJ_REV_STEP (one
 * launch per time step over all cells), MKH_ADJ_REV_BLOCK (one launch per
 * segment, a workgroup per domain), MKH_ADJ_REV_SHEET (one cooperative
 * launch per segment over the sheet's tiles) or MKH_ADJ_REV_STEP_OVERLAP
 * (as MKH_ADJ_REV_STEP, with each earlier segment recomputed on a second
 * stream while the later one's reverse steps run: small domains, whose
 * kernels leave most of the device idle), -1 before the first run; and
 * the number of reverse launches it made. No counterpart in the reference. */
#define MKH_ADJ_REV_STEP 0
#define MKH_ADJ_REV_BLOCK 1
#define MKH_ADJ_REV_SHEET 2
#define MKH_ADJ_REV_STEP_OVERLAP 3
int mkh_adj_reverse(void* sim, int* kind, long long* launches);

/* The caller's cell order for mkh_adj_run's field gradients: caller cell c
 * is device cell order[c] (0 <= order[c] < n_cells, n_cells = the
 * simulation's cells); the gradients are gathered into that order on the
 * device before they are copied out. order NULL or n_cells 0: the device's
 * order is the caller's. No counterpart in the reference, whose cells keep
 * the caller's order; for callers that renumber the cells of a connection
 * graph for locality. */
int mkh_adj_set_cell_order(void* sim, const long long* order,
                           long long n_cells);

/* The forward quantity of mkh_adj_run's loss: V at the probe cells after
 * each of the N time steps of tmin..tmax, into trace[s * n_probes + p]
 * (n_rows >= N). Leaves the state unchanged. No counterpart in the
 * reference; needs a kernel generated with the adjoint. */
int mkh_adj_trace(void* sim, double tmin, double tmax, double dt,
                  const double* events, long long n_events,
                  const long long* probes, long long n_probes, double* trace,
                  long long n_rows);

/* mkh_adj_trace without any device allocation. No counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc). Same
 * arguments, result and errors as mkh_adj_trace; the step table, the state
 * copies the run advances, the probe list and the device trace live in a
 * static block of MKH_STATIC_SCRATCH_BYTES per device (held until the
 * process ends) instead of the simulation's work buffers. A run whose
 * buffers exceed it (about 2 * n_cells * n_state * precision/8 bytes for
 * the states, plus 16-32 bytes per time step and 8 per probe and step)
 * fails with MKH_E_ARG before anything is launched. Every call on a device
 * uses the same block: calls that overlap on one device (other threads,
 * other simulations) race and must be serialised by the caller. The call
 * allocates no device or page-locked memory and frees none. */
#define MKH_STATIC_SCRATCH_BYTES (32LL << 20)
int mkh_adj_trace_static(void* sim, double tmin, double tmax, double dt,
                         const double* events, long long n_events,
                         const long long* probes, long long n_probes,
                         double* trace, long long n_rows);

/* The execution driver (MKH_DRIVER_*) the simulation currently uses. */
int mkh_sim_driver(void* sim, int* driver);

/* Number of log rows produced so far in the current run. */
int mkh_sim_log_rows(void* sim, long long* rows);

/* Host-only checks of the pacing system and the time-step planner (no
 * device needed). Counterpart of the reference's test harness
 * myokit/tests/ansic_event_based_pacing.c, which exposes ESys to Python.
 *   mkh_pacing_levels: creates the event system at t0, then for each
 *     times[k] (non-decreasing) advances and stores the level and the next
 *     event time.
 *   mkh_plan: the step sequence of a run (engine time, step size, pacing
 *     level, log flag per step), at most max_steps steps. */
int mkh_pacing_levels(const double* events, long long n_events, double t0,
                      const double* times, long long n_times, double* levels,
                      double* next_times);
int mkh_plan(double tmin, double tmax, double dt, const double* events,
             long long n_events, double log_interval, double* t,
             double* step, double* pace, int* log, long long max_steps,
             long long* n_steps);

#ifdef __cplusplus
}
#endif

#endif /* MYOKIT_AMD_H */
