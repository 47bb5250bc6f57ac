/*
 * mkh_host.cpp -- host side of the MI355X tissue engine (C ABI in
 * include/myokit_amd.h).
 *
 * Replaces the generated OpenCL host module of myokit/_sim/openclsim.c:
 *   - the kernel program is compiled once per generated source with hiprtc
 *     for the device's gfx target and cached (instead of clBuildProgram on
 *     every run, openclsim.c:815-839);
 *   - the time-step rule of sim_step (openclsim.c:1051-1162) runs on the
 *     host as a planner that writes a step table (time, dt, pace, log row)
 *     for a whole batch of steps; the device then runs the batch without
 *     host round trips, instead of 2-3 kernel enqueues plus blocking reads
 *     per step;
 *   - state, fields and logs stay resident in device memory (SoA), the log
 *     is copied back once per batch in bulk.
 */
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <numeric>
#include <unistd.h>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/myokit_amd.h"
#include "mk_args.h"
#include "mk_esys.h"

#include <mutex>

/* Device scratch of mkh_adj_trace_static, the entry point that allocates
 * nothing: one block per device, held from the library's load until the
 * process ends. On the host this name stands for the block but is not its
 * device address: host code uses only the address scratch_address() looks
 * up (per device, with that device current). */
static __device__ __align__(256) char scratch[MKH_STATIC_SCRATCH_BYTES];

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define MKH_HIP(call)                                                        \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess)                                                \
            return fail(MKH_E_HIP, std::string(#call) + " failed: " +        \
                                       hipGetErrorString(e_));               \
    } while (0)

/* -------------------------------------------------------------------------
 * Program: compiled kernel module
 * ---------------------------------------------------------------------- */
struct Program {
    int device = -1;
    hipModule_t module = nullptr;
    hipFunction_t f_cells = nullptr, f_block = nullptr, f_grid = nullptr;
    hipFunction_t f_sheet = nullptr, f_junction = nullptr;
    hipFunction_t f_chain = nullptr;
    hipFunction_t f_adj_step = nullptr, f_adj_block = nullptr,
                  f_adj_probe = nullptr, f_adj_sheet = nullptr,
                  f_adj_gather = nullptr;
    /* mk_layout of the module: precision, states, fields, sheet / chain /
     * reverse-sheet cells per thread (have_layout false if absent) */
    int layout[6] = {0, 0, 0, 0, 0, 0};
    bool have_layout = false;
    std::vector<char> code;
};

unsigned long long fnv1a(const std::string& s, unsigned long long h)
{
    for (unsigned char c : s) {
        h ^= c;
        h *= 1099511628211ULL;
    }
    return h;
}

std::vector<std::string> split_options(const char* options)
{
    std::vector<std::string> out;
    if (!options) return out;
    std::istringstream is(options);
    std::string w;
    while (is >> w) out.push_back(w);
    return out;
}

/* -------------------------------------------------------------------------
 * Time-step planner: the step and log rule of sim_step
 * (openclsim.c:1051-1162), with dt_min = 0.
 * ---------------------------------------------------------------------- */
struct StepRec {
    double t, dt, pace, t_after;
    bool log;
};

struct Planner {
    double tmin = 0, tmax = 0, dt0 = 0, log_interval = 0;
    double t = 0, tnext_log = 0, tnext_pace = 0, pace = 0;
    unsigned long long istep = 1, inext_log = 0;
    bool done = false;
    MkESys esys;

    int init(double tmin_, double tmax_, double dt_, const double* events,
             long long n_events, double log_interval_)
    {
        tmin = tmin_;
        tmax = tmax_;
        dt0 = dt_;
        log_interval = log_interval_;
        int f = esys.init(tmin, events, n_events);
        if (f != MkESys::OK) return f;
        tnext_pace = esys.next_time();
        pace = esys.level();
        t = tmin;
        istep = 1;
        inext_log = 0;
        tnext_log = tmin;
        done = !(tmax > tmin);
        return 0;
    }

    int next(StepRec& r)
    {
        const bool logging = t >= tnext_log;
        bool inter = false;
        double dt = tmin + (double)istep * dt0 - t;
        double d = tmax - t;
        if (d > 0 && d < dt) { dt = d; inter = true; }
        d = tnext_pace - t;
        if (d > 0 && d < dt) { dt = d; inter = true; }
        d = tnext_log - t;
        if (d > 0 && d < dt) { dt = d; inter = true; }
        if (!inter) istep++;
        r.t = t;
        r.dt = dt;
        r.pace = pace;
        r.log = logging;
        if (logging) {
            inext_log++;
            tnext_log = tmin + (double)inext_log * log_interval;
        }
        t += dt;
        r.t_after = t;
        int f = esys.advance(t);
        if (f != MkESys::OK) return f;
        tnext_pace = esys.next_time();
        pace = esys.level();
        if (t >= tmax) done = true;
        return 0;
    }
};

int pacing_error(int f)
{
    if (f == MkESys::SIMULTANEOUS)
        return fail(MKH_E_PACING_SIMULTANEOUS,
                    "E-Pacing error: Event scheduled or re-occuring at the "
                    "same time as another event.");
    if (f == MkESys::NEGATIVE_TIME)
        return fail(MKH_E_PACING_NEGATIVE_TIME,
                    "E-Pacing error: New time is before current time.");
    return fail(MKH_E_PACING_INVALID, "E-Pacing error: invalid protocol.");
}

/* Device address of `scratch` on each device, looked up on the first
 * call for that device (hipGetSymbolAddress with the device current) and
 * kept: each device holds its own copy of the block at its own address. */
std::mutex g_scratch_mu;
void* g_scratch_ptr[256] = {};

int scratch_address(int device, void** ptr)
{
    if (device < 0 || device >= 256)
        return fail(MKH_E_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int old = 0;
        MKH_HIP(hipGetDevice(&old));
        MKH_HIP(hipSetDevice(device));
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        (void)hipSetDevice(old);
        if (e != hipSuccess || !p)
            return fail(MKH_E_HIP, std::string("hipGetSymbolAddress(scratch) "
                                               "failed: ") +
                                       hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return 0;
}

/* Sets a driver variable for the lifetime of the guard. */
struct DriverAs {
    int& d;
    int saved;
    DriverAs(int& d_, int as) : d(d_), saved(d_) { d = as; }
    ~DriverAs() { d = saved; }
};

/* -------------------------------------------------------------------------
 * Simulation
 * ---------------------------------------------------------------------- */
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count)
    {
        if (count == n && p) return hipSuccess;
        release();
        if (count == 0) return hipSuccess;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e == hipSuccess) n = count;
        return e;
    }
};

struct SimBase {
    virtual ~SimBase() {}
    mkh_sim_config cfg{};
    Program* prog = nullptr;
    virtual int set_state(const double*, long long) = 0;
    virtual int get_state(double*, long long) = 0;
    virtual int set_fields(const double*, long long) = 0;
    virtual int set_conductance(double, double) = 0;
    virtual int set_conductance_field(const double*, long long, const double*,
                                      long long) = 0;
    virtual int set_connections(const long long*, const long long*,
                                const double*, long long) = 0;
    virtual int set_domains(const long long*, long long) = 0;
    virtual int set_pacing(int mode, const long long* a, long long n) = 0;
    virtual int set_logging(const int*, long long, int, int) = 0;
    virtual int init(double, double, double, const double*, long long, double,
                     long long*) = 0;
    virtual int set_log(void*, double*, double*) = 0;
    virtual int step(double*, int*) = 0;
    virtual int clean() = 0;
    virtual int set_option(const std::string&, long long) = 0;
    virtual int timing(double*, long long*, long long*) = 0;
    virtual int reset_timing() = 0;
    virtual long long log_rows() = 0;
    virtual int get_driver() = 0;
    virtual int adj_run(double, double, double, const double*, long long,
                        const long long*, long long, const double*, long long,
                        int, double*, double*, double*, double*) = 0;
    virtual int adj_timing(double*) = 0;
    virtual int adj_reverse(int*, long long*) = 0;
    virtual int set_cell_order(const long long*, long long) = 0;
    virtual int copy_state_from(SimBase* src) = 0;
    virtual int set_nan_cell(long long c) = 0;
    virtual int trace_static(double, double, double, const double*,
                             long long, const long long*, long long, double*,
                             long long) = 0;
};

template <typename R>
struct Sim : SimBase {
    typedef MkArgsT<R> Args;
    typedef MkStepT<R> Step;

    long long n = 0, npad = 0;
    int cur = 0;                  /* state buffer holding the current state */
    hipStream_t stream = nullptr;
    DevBuf<R> state[2], fields, csr_g, gxf, gyf, logbuf;
    DevBuf<int> colmap, csr_ptr, csr_idx, dom_cells, halt;
    DevBuf<long long> dom_off;
    DevBuf<unsigned char> pace_mask;
    DevBuf<Step> steps;
    R gx = 0, gy = 0;
    int pace_mode = 0;
    long long px0 = 0, px1 = 0, py0 = 0, py1 = 0;
    long long nan_cell = 0;       /* device index of the caller's cell 0 */
    int ncols = 0, nan_check = 0;
    long long nquant = 0;
    int driver = MKH_DRIVER_GRID;
    long long ndom = 1;
    /* sheet driver */
    /* tile edges as step-tagged words: one per 32 bits of a value */
    static constexpr int XW = (int)(sizeof(R) / 4);
    DevBuf<unsigned long long> sh_edges;
    DevBuf<int> sh_err;
    /* chain driver (graphs whose every connection joins cells c, c + 1) */
    DevBuf<R> ch_gl, ch_gr;
    DevBuf<unsigned long long> ch_edges;
    DevBuf<unsigned char> ch_cf;
    DevBuf<int> ch_err;
    bool is_chain = false;
    int ch_tiles = 0, ch_len = 0;
    long long ch_steps = 0;
    int sh_tx = 0, sh_ty = 0, sh_capacity = 0;
    long long sh_steps = 0;
    long long max_domain = 0;
    int max_block_cells = 0;

    /* graph connections kept on the host for re-layout */
    std::vector<long long> ci, cj;
    std::vector<double> cg;
    bool have_graph = false;
    /* graphs that are the caller's identical copies, interleaved (device
     * cell p * copies + k is copy k's cell p): copies asked for, and the
     * copies of the CSR in use (1: one row per cell) */
    long long graph_copies = 1;
    int csr_copies = 1;

    /* options */
    long long batch_steps = 1 << 16;
    long long log_rows_cap = 0;   /* 0 = automatic */
    bool time_kernels = false;
    int forced_driver = -1;

    /* run state */
    bool running = false;
    Planner plan;
    long long total_rows = 0, rows_done = 0;
    R* dst_values = nullptr;
    double* dst_times = nullptr;
    double* dst_paces = nullptr;
    Step* h_steps = nullptr;       /* pinned staging */
    size_t h_steps_cap = 0;
    std::vector<double> row_t_after;

    /* timing */
    double kernel_ms = 0;
    long long launches = 0, cell_steps = 0;
    std::vector<hipEvent_t> ev_pool;

    /* Log rows go to the host on a second stream (default priority), each
     * as soon as the launch that wrote it has finished, while the next
     * launches run. row_ev is a small ring of "row ready" events. */
    hipStream_t copy_stream = nullptr;
    hipEvent_t row_ev[16] = {};

    ~Sim() override
    {
        if (h_steps) (void)hipHostFree(h_steps);
        for (auto e : ev_pool) (void)hipEventDestroy(e);
        for (auto e : row_ev)
            if (e) (void)hipEventDestroy(e);
        for (auto e : rc_ev)
            if (e) (void)hipEventDestroy(e);
        if (rc_stream) (void)hipStreamDestroy(rc_stream);
        if (copy_stream) (void)hipStreamDestroy(copy_stream);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int create(Program* p, const mkh_sim_config* c)
    {
        prog = p;
        cfg = *c;
        n = cfg.nx * cfg.ny;
        npad = (n + 255) / 256 * 256;
        max_block_cells = cfg.precision == 64 ? 2048 : 4096;
        MKH_HIP(hipSetDevice(cfg.device));
        MKH_HIP(hipStreamCreate(&stream));
        MKH_HIP(hipStreamCreate(&copy_stream));
        for (auto& e : row_ev)
            MKH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        MKH_HIP(state[0].alloc((size_t)cfg.n_state * npad));
        MKH_HIP(state[1].alloc((size_t)cfg.n_state * npad));
        MKH_HIP(hipMemset(state[0].p, 0, state[0].n * sizeof(R)));
        MKH_HIP(hipMemset(state[1].p, 0, state[1].n * sizeof(R)));
        if (cfg.n_field > 0) {
            MKH_HIP(fields.alloc((size_t)cfg.n_field * npad));
            MKH_HIP(hipMemset(fields.p, 0, fields.n * sizeof(R)));
        }
        MKH_HIP(halt.alloc(1));
        int m1 = -1;
        MKH_HIP(hipMemcpy(halt.p, &m1, sizeof(int), hipMemcpyHostToDevice));
        long long off[2] = {0, n};
        MKH_HIP(dom_off.alloc(2));
        MKH_HIP(hipMemcpy(dom_off.p, off, sizeof(off), hipMemcpyHostToDevice));
        ndom = 1;
        max_domain = n;
        pace_mode = 0;
        if (cfg.sheet_cpt > 0 && cfg.ny > 1 && p->f_sheet) {
            int per_cu = 0, dev = cfg.device;
            MKH_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, p->f_sheet, 256, 0));
            hipDeviceProp_t prop;
            MKH_HIP(hipGetDeviceProperties(&prop, dev));
            sh_capacity = per_cu * prop.multiProcessorCount;
            const int th = 4 * cfg.sheet_cpt;
            sh_tx = (int)((cfg.nx + 63) / 64);
            sh_ty = (int)((cfg.ny + th - 1) / th);
            if ((long long)sh_tx * sh_ty <= sh_capacity) {
                const size_t nt = (size_t)sh_tx * sh_ty;
                MKH_HIP(sh_edges.alloc(2 * nt * 4 * 64 * XW));
                MKH_HIP(sh_err.alloc(1));
                MKH_HIP(hipMemset(sh_edges.p, 0,
                                  sh_edges.n * sizeof(unsigned long long)));
                MKH_HIP(hipMemset(sh_err.p, 0, sizeof(int)));
            }
        }
        choose_driver();
        return 0;
    }

    /* Whether driver d can run this geometry and coupling: the block
     * driver needs every domain in one workgroup's LDS, the sheet a 2D
     * homogeneous stencil whose tiles are co-resident, the cells driver
     * uncoupled cells. */
    bool feasible(int d) const
    {
        switch (d) {
        case MKH_DRIVER_CELLS:
            return cfg.coupling == MKH_COUPLING_NONE;
        case MKH_DRIVER_BLOCK:
            return max_domain <= max_block_cells;
        case MKH_DRIVER_GRID:
            return true;
        case MKH_DRIVER_CHAIN:
            return cfg.coupling == MKH_COUPLING_GRAPH && is_chain &&
                   ch_tiles > 0 && ch_edges.p;
        case MKH_DRIVER_SHEET:
            return cfg.coupling == MKH_COUPLING_STENCIL && cfg.ny > 1 &&
                   sh_tx > 0 && sh_ty > 0 && sh_edges.p &&
                   (long long)sh_tx * sh_ty <= sh_capacity;
        default:
            return false;
        }
    }

    void choose_driver()
    {
        if (forced_driver >= 0 && feasible(forced_driver)) {
            driver = forced_driver;
            return;
        }
        if (cfg.coupling == MKH_COUPLING_NONE) {
            driver = MKH_DRIVER_CELLS;
        } else if (cfg.coupling == MKH_COUPLING_STENCIL && cfg.ny > 1 &&
                   n > 4096 && sh_tx > 0 &&
                   (long long)sh_tx * sh_ty <= sh_capacity) {
            driver = MKH_DRIVER_SHEET;
        } else if (max_domain <= max_block_cells &&
                   (ndom >= 64 || max_domain <= 512)) {
            driver = MKH_DRIVER_BLOCK;
        } else {
            driver = MKH_DRIVER_GRID;
        }
    }

    int set_state(const double* s, long long count) override
    {
        if (count != n * cfg.n_state)
            return fail(MKH_E_ARG, "state has the wrong size");
        std::vector<R> soa((size_t)cfg.n_state * npad, (R)0);
        for (long long c = 0; c < n; ++c)
            for (int i = 0; i < cfg.n_state; ++i)
                soa[(size_t)i * npad + c] = (R)s[c * cfg.n_state + i];
        MKH_HIP(hipMemcpy(state[cur].p, soa.data(), soa.size() * sizeof(R),
                          hipMemcpyHostToDevice));
        return 0;
    }

    /* Device-to-device copy of another simulation's current state (same
     * precision, state count and cell count, so the same SoA layout). */
    int copy_state_from(SimBase* src_base) override
    {
        Sim<R>* src = dynamic_cast<Sim<R>*>(src_base);
        if (!src || src->cfg.n_state != cfg.n_state || src->n != n ||
            src->npad != npad)
            return fail(MKH_E_ARG, "states of different layouts");
        MKH_HIP(hipStreamSynchronize(src->stream));
        MKH_HIP(hipMemcpyAsync(state[cur].p, src->state[src->cur].p,
                               (size_t)cfg.n_state * npad * sizeof(R),
                               hipMemcpyDeviceToDevice, stream));
        MKH_HIP(hipStreamSynchronize(stream));
        return 0;
    }

    int get_state(double* s, long long count) override
    {
        if (count != n * cfg.n_state)
            return fail(MKH_E_ARG, "state has the wrong size");
        std::vector<R> soa((size_t)cfg.n_state * npad);
        MKH_HIP(hipStreamSynchronize(stream));
        MKH_HIP(hipMemcpy(soa.data(), state[cur].p, soa.size() * sizeof(R),
                          hipMemcpyDeviceToHost));
        for (long long c = 0; c < n; ++c)
            for (int i = 0; i < cfg.n_state; ++i)
                s[c * cfg.n_state + i] = (double)soa[(size_t)i * npad + c];
        return 0;
    }

    int set_fields(const double* d, long long count) override
    {
        if (count != n * cfg.n_field)
            return fail(MKH_E_ARG, "field data has the wrong size");
        if (cfg.n_field == 0) return 0;
        std::vector<R> soa((size_t)cfg.n_field * npad, (R)0);
        for (long long c = 0; c < n; ++c)
            for (int k = 0; k < cfg.n_field; ++k)
                soa[(size_t)k * npad + c] = (R)d[c * cfg.n_field + k];
        MKH_HIP(hipMemcpy(fields.p, soa.data(), soa.size() * sizeof(R),
                          hipMemcpyHostToDevice));
        return 0;
    }

    int set_conductance(double x, double y) override
    {
        gx = (R)x;
        gy = (R)y;
        return 0;
    }

    int set_conductance_field(const double* x, long long nxf,
                              const double* y, long long nyf) override
    {
        if (nxf != (cfg.nx - 1) * cfg.ny)
            return fail(MKH_E_ARG, "gx field has the wrong size");
        if (cfg.ny > 1 && nyf != (cfg.ny - 1) * cfg.nx)
            return fail(MKH_E_ARG, "gy field has the wrong size");
        std::vector<R> hx(std::max<long long>(nxf, 1), (R)0);
        for (long long k = 0; k < nxf; ++k) hx[k] = (R)x[k];
        MKH_HIP(gxf.alloc(hx.size()));
        MKH_HIP(hipMemcpy(gxf.p, hx.data(), hx.size() * sizeof(R),
                          hipMemcpyHostToDevice));
        std::vector<R> hy(std::max<long long>(nyf, 1), (R)0);
        for (long long k = 0; k < nyf && y; ++k) hy[k] = (R)y[k];
        MKH_HIP(gyf.alloc(hy.size()));
        MKH_HIP(hipMemcpy(gyf.p, hy.data(), hy.size() * sizeof(R),
                          hipMemcpyHostToDevice));
        return 0;
    }

    int set_connections(const long long* i, const long long* j,
                        const double* g, long long m) override
    {
        ci.assign(i, i + m);
        cj.assign(j, j + m);
        cg.assign(g, g + m);
        for (long long k = 0; k < m; ++k) {
            if (ci[k] < 0 || ci[k] >= n || cj[k] < 0 || cj[k] >= n ||
                ci[k] == cj[k])
                return fail(MKH_E_ARG, "invalid connection");
        }
        have_graph = true;
        /* Connected components (union-find) */
        std::vector<long long> parent(n);
        std::iota(parent.begin(), parent.end(), 0);
        auto find = [&](long long a) {
            while (parent[a] != a) {
                parent[a] = parent[parent[a]];
                a = parent[a];
            }
            return a;
        };
        for (long long k = 0; k < m; ++k) {
            long long a = find(ci[k]), b = find(cj[k]);
            if (a != b) parent[std::max(a, b)] = std::min(a, b);
        }
        std::vector<long long> root(n), comp_size(n, 0);
        for (long long c = 0; c < n; ++c) {
            root[c] = find(c);
            comp_size[root[c]]++;
        }
        max_domain = 0;
        for (long long c = 0; c < n; ++c)
            max_domain = std::max(max_domain, comp_size[c]);
        int e = build_chain_layout();
        if (e) return e;
        if (forced_driver >= 0 && forced_driver != MKH_DRIVER_GRID &&
            !feasible(forced_driver))
            return fail(MKH_E_ARG, "driver " + std::to_string(forced_driver) +
                                       " cannot run this geometry and "
                                       "coupling");
        choose_driver_graph();
        return build_graph_layout(root, comp_size);
    }

    /* The chain layout: used when every connection joins cells c and c + 1
     * (each at most once), i.e. the graph is a set of disjoint chains laid
     * end to end in the device's cell order. Per cell: the conductances to
     * c - 1 and c + 1, which of them exist, and which comes first in the
     * cell's connection-list order (the order of its sum). The cells are cut
     * into one-wave tiles of equal length (the last one shorter). */
    int build_chain_layout()
    {
        is_chain = false;
        ch_tiles = ch_len = 0;
        const long long m = (long long)ci.size();
        if (m == 0 || !prog->f_chain || cfg.chain_cpt < 1) return 0;
        std::vector<R> gl(npad, (R)0), gr(npad, (R)0);
        std::vector<unsigned char> cf(npad, 0);
        for (long long k = 0; k < m; ++k) {
            const long long a = std::min(ci[k], cj[k]);
            const long long b = std::max(ci[k], cj[k]);
            if (b != a + 1 || (cf[a] & 2u)) return 0;
            gr[a] = (R)cg[k];
            gl[b] = (R)cg[k];
            /* the first connection listed for a cell sets its order */
            if (!(cf[a] & 1u)) cf[a] &= (unsigned char)~4u;
            if (!(cf[b] & 2u)) cf[b] |= 4u;
            cf[a] |= 2u;
            cf[b] |= 1u;
        }
        int per_cu = 0, ncu = 0;
        MKH_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, prog->f_chain, 64, 0));
        MKH_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                                      cfg.device));
        const long long cap = (long long)per_cu * ncu;
        /* 64 * MK_CCPT cells per tile, fewer if the tiles would then not
         * give every SIMD a wave; more tiles than can be co-resident: as
         * many equal tiles as can */
        const long long want = 64LL * cfg.chain_cpt;
        const long long simds = 4LL * ncu;
        long long len = std::min<long long>(
            want, std::max<long long>(64, ((n + simds - 1) / simds + 63) / 64
                                               * 64));
        if ((n + len - 1) / len > cap) {
            if (cap < 1) return 0;
            len = (n + cap - 1) / cap;
        }
        if (len > want) return 0;
        const long long tiles = (n + len - 1) / len;
        MKH_HIP(ch_gl.alloc(npad));
        MKH_HIP(ch_gr.alloc(npad));
        MKH_HIP(ch_cf.alloc(npad));
        MKH_HIP(hipMemcpy(ch_gl.p, gl.data(), npad * sizeof(R),
                          hipMemcpyHostToDevice));
        MKH_HIP(hipMemcpy(ch_gr.p, gr.data(), npad * sizeof(R),
                          hipMemcpyHostToDevice));
        MKH_HIP(hipMemcpy(ch_cf.p, cf.data(), npad, hipMemcpyHostToDevice));
        MKH_HIP(ch_edges.alloc((size_t)tiles * 4 * XW));
        MKH_HIP(ch_err.alloc(1));
        MKH_HIP(hipMemset(ch_edges.p, 0,
                          ch_edges.n * sizeof(unsigned long long)));
        MKH_HIP(hipMemset(ch_err.p, 0, sizeof(int)));
        ch_steps = 0;
        ch_tiles = (int)tiles;
        ch_len = (int)len;
        is_chain = true;
        return 0;
    }

    void choose_driver_graph()
    {
        ndom = 0;   /* set by build_graph_layout */
        if (forced_driver == MKH_DRIVER_GRID ||
            (forced_driver == MKH_DRIVER_BLOCK &&
             feasible(MKH_DRIVER_BLOCK)) ||
            (forced_driver == MKH_DRIVER_CHAIN &&
             feasible(MKH_DRIVER_CHAIN))) {
            driver = forced_driver;
            return;
        }
        if (feasible(MKH_DRIVER_CHAIN) && forced_driver < 0)
            driver = MKH_DRIVER_CHAIN;
        else
            driver = max_domain <= max_block_cells ? MKH_DRIVER_BLOCK
                                                   : MKH_DRIVER_GRID;
    }

    /* The CSR of one copy, for a graph of graph_copies identical copies
     * held interleaved: the connection list is the copies' lists one after
     * the other, copy k's connection e joining cells (a + k, b + k) with
     * the conductance of copy 0's connection e (a, b multiples of the
     * number of copies). Rows are copy 0's cells p = a / copies, entries in
     * list order with neighbour positions; a kernel maps device cell c to
     * row c / copies and neighbour q to q * copies + c % copies, so every
     * cell sums its own connections in its own list order, as with a row
     * per cell. Returns false (and uploads nothing) if the graph is not
     * such a set of copies. */
    bool copies_layout()
    {
        const long long K = graph_copies;
        const long long m = (long long)ci.size();
        if (K < 2 || n % K || m % K || m == 0 || n / K > INT32_MAX)
            return false;
        const long long E = m / K, nc = n / K;
        for (long long e = 0; e < E; ++e)
            if (ci[e] % K || cj[e] % K) return false;
        for (long long k = 1; k < K; ++k)
            for (long long e = 0; e < E; ++e) {
                const long long f = k * E + e;
                if (ci[f] != ci[e] + k || cj[f] != cj[e] + k ||
                    cg[f] != cg[e])
                    return false;
            }
        std::vector<int> cnt(nc + 1, 0);
        for (long long e = 0; e < E; ++e) {
            cnt[ci[e] / K + 1]++;
            cnt[cj[e] / K + 1]++;
        }
        for (long long p = 0; p < nc; ++p) cnt[p + 1] += cnt[p];
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        std::vector<int> hidx(std::max<long long>(2 * E, 1), 0);
        std::vector<R> hg(std::max<long long>(2 * E, 1), (R)0);
        for (long long e = 0; e < E; ++e) {
            const int pa = (int)(ci[e] / K), pb = (int)(cj[e] / K);
            hidx[fill[pa]] = pb;
            hg[fill[pa]++] = (R)cg[e];
            hidx[fill[pb]] = ~pa;
            hg[fill[pb]++] = (R)cg[e];
        }
        if (csr_ptr.alloc(cnt.size()) != hipSuccess ||
            hipMemcpy(csr_ptr.p, cnt.data(), cnt.size() * 4,
                      hipMemcpyHostToDevice) != hipSuccess ||
            csr_idx.alloc(hidx.size()) != hipSuccess ||
            hipMemcpy(csr_idx.p, hidx.data(), hidx.size() * 4,
                      hipMemcpyHostToDevice) != hipSuccess ||
            csr_g.alloc(hg.size()) != hipSuccess ||
            hipMemcpy(csr_g.p, hg.data(), hg.size() * sizeof(R),
                      hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        csr_copies = (int)K;
        ndom = 1;
        return true;
    }

    int build_graph_layout(const std::vector<long long>& root,
                           const std::vector<long long>& comp_size)
    {
        const long long m = (long long)ci.size();
        csr_copies = 1;
        if (driver == MKH_DRIVER_BLOCK) {
            /* Domains = components, cells in increasing id order; CSR rows in
             * domain order with neighbour indices local to the domain. */
            std::vector<long long> order(n);
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(),
                             [&](long long a, long long b) {
                                 return root[a] < root[b];
                             });
            std::vector<long long> pos(n), offs;
            std::vector<long long> dstart(n, -1);
            for (long long p = 0; p < n; ++p) {
                long long c = order[p];
                if (p == 0 || root[order[p - 1]] != root[c]) {
                    offs.push_back(p);
                    dstart[root[c]] = p;
                }
                pos[c] = p;
            }
            offs.push_back(n);
            ndom = (long long)offs.size() - 1;
            std::vector<std::vector<std::pair<int, R>>> rows(n);
            for (long long k = 0; k < m; ++k) {
                long long a = ci[k], b = cj[k];
                long long base = dstart[root[a]];
                rows[pos[a]].push_back({(int)(pos[b] - base), (R)cg[k]});
                rows[pos[b]].push_back({~(int)(pos[a] - base), (R)cg[k]});
            }
            std::vector<int> hptr(n + 1, 0), hidx, hcells(n);
            std::vector<R> hg;
            for (long long p = 0; p < n; ++p) {
                hptr[p + 1] = hptr[p] + (int)rows[p].size();
                for (auto& e : rows[p]) {
                    hidx.push_back(e.first);
                    hg.push_back(e.second);
                }
                hcells[p] = (int)order[p];
            }
            if (hidx.empty()) {
                hidx.push_back(0);
                hg.push_back((R)0);
            }
            MKH_HIP(csr_ptr.alloc(hptr.size()));
            MKH_HIP(hipMemcpy(csr_ptr.p, hptr.data(), hptr.size() * 4,
                              hipMemcpyHostToDevice));
            MKH_HIP(csr_idx.alloc(hidx.size()));
            MKH_HIP(hipMemcpy(csr_idx.p, hidx.data(), hidx.size() * 4,
                              hipMemcpyHostToDevice));
            MKH_HIP(csr_g.alloc(hg.size()));
            MKH_HIP(hipMemcpy(csr_g.p, hg.data(), hg.size() * sizeof(R),
                              hipMemcpyHostToDevice));
            MKH_HIP(dom_cells.alloc(hcells.size()));
            MKH_HIP(hipMemcpy(dom_cells.p, hcells.data(), hcells.size() * 4,
                              hipMemcpyHostToDevice));
            MKH_HIP(dom_off.alloc(offs.size()));
            MKH_HIP(hipMemcpy(dom_off.p, offs.data(),
                              offs.size() * sizeof(long long),
                              hipMemcpyHostToDevice));
        } else if (copies_layout()) {
            /* done: one copy's CSR, shared by all copies */
        } else {
            /* Global CSR by cell id, entries in connection-list order. */
            std::vector<int> cnt(n + 1, 0);
            for (long long k = 0; k < m; ++k) {
                cnt[ci[k] + 1]++;
                cnt[cj[k] + 1]++;
            }
            for (long long c = 0; c < n; ++c) cnt[c + 1] += cnt[c];
            std::vector<int> fill(cnt.begin(), cnt.end() - 1);
            std::vector<int> hidx(std::max<long long>(2 * m, 1), 0);
            std::vector<R> hg(std::max<long long>(2 * m, 1), (R)0);
            for (long long k = 0; k < m; ++k) {
                int a = (int)ci[k], b = (int)cj[k];
                hidx[fill[a]] = b;
                hg[fill[a]++] = (R)cg[k];
                hidx[fill[b]] = ~a;
                hg[fill[b]++] = (R)cg[k];
            }
            MKH_HIP(csr_ptr.alloc(cnt.size()));
            MKH_HIP(hipMemcpy(csr_ptr.p, cnt.data(), cnt.size() * 4,
                              hipMemcpyHostToDevice));
            MKH_HIP(csr_idx.alloc(hidx.size()));
            MKH_HIP(hipMemcpy(csr_idx.p, hidx.data(), hidx.size() * 4,
                              hipMemcpyHostToDevice));
            MKH_HIP(csr_g.alloc(hg.size()));
            MKH_HIP(hipMemcpy(csr_g.p, hg.data(), hg.size() * sizeof(R),
                              hipMemcpyHostToDevice));
            ndom = 1;
        }
        (void)comp_size;
        return 0;
    }

    int set_domains(const long long* offs, long long nd) override
    {
        if (nd < 1 || offs[0] != 0 || offs[nd] != n)
            return fail(MKH_E_ARG, "domain offsets must run from 0 to n");
        max_domain = 0;
        for (long long d = 0; d < nd; ++d) {
            if (offs[d + 1] < offs[d])
                return fail(MKH_E_ARG, "domain offsets must increase");
            max_domain = std::max(max_domain, offs[d + 1] - offs[d]);
        }
        MKH_HIP(dom_off.alloc(nd + 1));
        MKH_HIP(hipMemcpy(dom_off.p, offs, (nd + 1) * sizeof(long long),
                          hipMemcpyHostToDevice));
        ndom = nd;
        dom_cells.release();
        choose_driver();
        return 0;
    }

    int set_pacing(int mode, const long long* a, long long m) override
    {
        pace_mode = mode;
        if (mode == 1) {
            /* rectangle (nx, ny, x, y) as normalised by set_paced_cells */
            px0 = a[2];
            px1 = a[2] + a[0];
            py0 = a[3];
            py1 = a[3] + a[1];
        } else if (mode == 2) {
            std::vector<unsigned char> mask(npad, 0);
            for (long long k = 0; k < m; ++k) {
                if (a[k] < 0 || a[k] >= n)
                    return fail(MKH_E_ARG, "paced cell out of range");
                mask[a[k]] = 1;
            }
            MKH_HIP(pace_mask.alloc(npad));
            MKH_HIP(hipMemcpy(pace_mask.p, mask.data(), npad,
                              hipMemcpyHostToDevice));
        }
        return 0;
    }

    /* the column map last uploaded (as passed in), so that repeated runs
     * with the same logging skip the upload */
    std::vector<int> cmap_last;

    int set_nan_cell(long long c) override
    {
        if (c < 0 || c >= n) return fail(MKH_E_ARG, "NaN cell out of range");
        nan_cell = c;
        return 0;
    }

    int set_logging(const int* cmap, long long nq, int nc, int nanc) override
    {
        if (nq != cfg.n_quant)
            return fail(MKH_E_ARG, "quantity count differs from the kernel");
        nquant = nq;
        ncols = nc;
        nan_check = nanc;
        if (nq > 0) {
            const size_t m = (size_t)nq * n;
            if (colmap.p && cmap_last.size() == m &&
                std::memcmp(cmap_last.data(), cmap, m * sizeof(int)) == 0)
                return 0;
            std::vector<int> h((size_t)nq * npad, -1);
            for (long long q = 0; q < nq; ++q)
                for (long long c = 0; c < n; ++c)
                    h[(size_t)q * npad + c] = cmap[q * n + c];
            cmap_last.clear();
            MKH_HIP(colmap.alloc(h.size()));
            MKH_HIP(hipMemcpy(colmap.p, h.data(), h.size() * sizeof(int),
                              hipMemcpyHostToDevice));
            cmap_last.assign(cmap, cmap + m);
        }
        return 0;
    }

    int init(double tmin, double tmax, double dt, const double* events,
             long long nev, double log_interval, long long* nrows) override
    {
        if (!(dt > 0)) return fail(MKH_E_ARG, "step size must be positive");
        if (!(log_interval > 0))
            return fail(MKH_E_ARG, "log interval must be positive");
        int f = plan.init(tmin, tmax, dt, events, nev, log_interval);
        if (f) return pacing_error(f);
        /* Dry run to count the log rows */
        Planner dry = plan;
        long long rows = 0;
        StepRec r;
        while (!dry.done) {
            f = dry.next(r);
            if (f) return pacing_error(f);
            if (r.log) rows++;
        }
        total_rows = rows;
        rows_done = 0;
        if (nrows) *nrows = rows;
        dst_values = nullptr;
        dst_times = dst_paces = nullptr;
        int m1 = -1;
        MKH_HIP(hipMemcpy(halt.p, &m1, sizeof(int), hipMemcpyHostToDevice));
        /* Device log buffer: whole run if it fits a budget, else batches */
        long long cap = log_rows_cap;
        if (cap <= 0) {
            const long long budget = 1LL << 30;  /* bytes */
            long long row_bytes = std::max<long long>(1, ncols) * sizeof(R);
            cap = std::max<long long>(1, budget / row_bytes);
        }
        cap = std::max<long long>(1, std::min<long long>(cap, rows));
        if (ncols > 0 && rows > 0) {
            MKH_HIP(logbuf.alloc((size_t)cap * ncols));
        }
        log_rows_cap_run = cap;
        /* the step tags restart: old tags must not pass for new ones */
        if (sh_edges.p) {
            MKH_HIP(hipMemset(sh_edges.p, 0,
                              sh_edges.n * sizeof(unsigned long long)));
            MKH_HIP(hipMemset(sh_err.p, 0, sizeof(int)));
            sh_steps = 0;
        }
        if (ch_edges.p) {
            MKH_HIP(hipMemset(ch_edges.p, 0,
                              ch_edges.n * sizeof(unsigned long long)));
            MKH_HIP(hipMemset(ch_err.p, 0, sizeof(int)));
            ch_steps = 0;
        }
        size_t need = (size_t)batch_steps;
        MKH_HIP(steps.alloc(need));
        if (h_steps_cap < need) {
            if (h_steps) (void)hipHostFree(h_steps);
            MKH_HIP(hipHostMalloc((void**)&h_steps, need * sizeof(Step)));
            h_steps_cap = need;
        }
        running = true;
        return 0;
    }
    long long log_rows_cap_run = 1;

    int set_log(void* v, double* t, double* p) override
    {
        dst_values = (R*)v;
        dst_times = t;
        dst_paces = p;
        return 0;
    }

    Args base_args()
    {
        Args a;
        std::memset(&a, 0, sizeof(a));
        a.state_a = state[cur].p;
        a.state_b = state[cur ^ 1].p;
        a.fields = fields.p;
        a.steps = steps.p;
        a.logbuf = logbuf.p;
        a.colmap = colmap.p;
        a.pace_mask = pace_mask.p;
        a.csr_ptr = csr_ptr.p;
        a.csr_idx = csr_idx.p;
        a.csr_g = csr_g.p;
        a.gcopies = csr_copies;
        a.gx_field = gxf.p;
        a.gy_field = gyf.p;
        a.dom_off = dom_off.p;
        a.dom_cells = dom_cells.p;
        a.halt = halt.p;
        a.n = n;
        a.npad = npad;
        a.nx = cfg.nx;
        a.ny = cfg.ny;
        a.px0 = px0;
        a.px1 = px1;
        a.py0 = py0;
        a.py1 = py1;
        a.pace_mode = pace_mode;
        a.ncols = ncols;
        a.nan_check = nan_check;
        a.nan_cell = nan_cell;
        a.ndom = (int)ndom;
        a.gx = gx;
        a.gy = gy;
        return a;
    }

    MkSheetT<R> sheet_args()
    {
        MkSheetT<R> sh;
        std::memset(&sh, 0, sizeof(sh));
        sh.edges = sh_edges.p;
        sh.err = sh_err.p;
        sh.tiles_x = sh_tx;
        sh.tiles_y = sh_ty;
        sh.spin_limit = 1 << 20;
        sh.traj = nullptr;
        return sh;
    }

    MkChainT<R> chain_args()
    {
        MkChainT<R> ch;
        std::memset(&ch, 0, sizeof(ch));
        ch.gl = ch_gl.p;
        ch.gr = ch_gr.p;
        ch.cf = ch_cf.p;
        ch.edges = ch_edges.p;
        ch.err = ch_err.p;
        ch.tiles = ch_tiles;
        ch.tile_len = ch_len;
        ch.spin_limit = 1 << 20;
        return ch;
    }

    hipEvent_t event(size_t k)
    {
        while (ev_pool.size() <= k) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            ev_pool.push_back(e);
        }
        return ev_pool[k];
    }

    int launch(hipFunction_t f, unsigned grid, void** params, size_t& nev)
    {
        const unsigned bt = (unsigned)cfg.block_threads;
        if (time_kernels) MKH_HIP(hipEventRecord(event(nev), stream));
        MKH_HIP(hipModuleLaunchKernel(f, grid, 1, 1, bt, 1, 1, 0, stream,
                                      params, nullptr));
        if (time_kernels) MKH_HIP(hipEventRecord(event(nev + 1), stream));
        if (time_kernels) nev += 2;
        launches++;
        return 0;
    }

    /* A batch of planned steps. step() plans, launches and finishes one
     * batch for this simulation; the fiber-tissue driver interleaves the
     * launches of two simulations step by step. */
    struct Batch {
        long long ns = 0, nr = 0;
        std::vector<double> tlog, plog;
        size_t nev = 0;
    };

    int plan_batch(Batch& b)
    {
        b = Batch();
        row_t_after.clear();
        StepRec r;
        while (!plan.done && b.ns < batch_steps && b.nr < log_rows_cap_run) {
            int f = plan.next(r);
            if (f) return pacing_error(f);
            Step& st = h_steps[b.ns++];
            st.t = (R)r.t;
            st.dt = (R)r.dt;
            st.pace = (R)r.pace;
            st.row = -1;
            if (r.log) {
                st.row = (int)b.nr++;
                row_t_after.push_back(r.t_after);
                b.tlog.push_back((double)(R)r.t);
                b.plog.push_back((double)(R)r.pace);
            }
        }
        if (b.ns > 0)
            MKH_HIP(hipMemcpyAsync(steps.p, h_steps, b.ns * sizeof(Step),
                                   hipMemcpyHostToDevice, stream));
        return 0;
    }

    /* After the launch that ends with step s: if s logs a row, copy the row
     * to the caller's buffer on the copy stream. */
    int emit_row(const Batch& b, long long s)
    {
        const int row = h_steps[s].row;
        if (!(b.nr > 0 && ncols > 0 && dst_values) || row < 0) return 0;
        hipEvent_t e = row_ev[row % 16];
        MKH_HIP(hipEventRecord(e, stream));
        MKH_HIP(hipStreamWaitEvent(copy_stream, e, 0));
        MKH_HIP(hipMemcpyAsync(
            dst_values + (size_t)(rows_done + row) * ncols,
            logbuf.p + (size_t)row * ncols, (size_t)ncols * sizeof(R),
            hipMemcpyDeviceToHost, copy_stream));
        return 0;
    }

    unsigned grid_blocks() const
    {
        const unsigned bt = (unsigned)cfg.block_threads;
        if ((cfg.coupling == MKH_COUPLING_STENCIL ||
             cfg.coupling == MKH_COUPLING_HETERO) && cfg.ny > 1) {
            long long tx = (cfg.nx + 63) / 64;
            long long ty = (cfg.ny + (bt / 64) - 1) / (bt / 64);
            return (unsigned)(tx * ty);
        }
        return (unsigned)((n + bt - 1) / bt);
    }

    /* Step s of the batch with the grid driver; `extra` (or null) is a
     * current per cell added to the diffusion current where nonzero. */
    int grid_launch(Batch& b, long long s, const R* extra)
    {
        Args a = base_args();
        a.idiff_extra = extra;
        R* src = state[cur].p;
        R* dst = state[cur ^ 1].p;
        int si = (int)s;
        void* params[] = {&a, &src, &dst, &si};
        int e = launch(prog->f_grid, grid_blocks(), params, b.nev);
        if (e) return e;
        cur ^= 1;
        return emit_row(b, s);
    }

    /* Waits for the batch; copies times and paces of its rows. Returns the
     * number of valid rows (all, or up to the one where the NaN check
     * halted the run) through `valid`. */
    int finish_batch(Batch& b, long long& valid, int& halted)
    {
        cell_steps += b.ns * n;
        int h = -1;
        MKH_HIP(hipMemcpyAsync(&h, halt.p, sizeof(int), hipMemcpyDeviceToHost,
                               stream));
        MKH_HIP(hipStreamSynchronize(stream));
        MKH_HIP(hipStreamSynchronize(copy_stream));
        if (driver == MKH_DRIVER_SHEET || driver == MKH_DRIVER_CHAIN) {
            int e = 0;
            MKH_HIP(hipMemcpy(&e, driver == MKH_DRIVER_SHEET ? sh_err.p
                                                             : ch_err.p,
                              sizeof(int), hipMemcpyDeviceToHost));
            if (e) {
                running = false;
                return fail(MKH_E_HIP, std::string(driver == MKH_DRIVER_SHEET
                                                       ? "sheet" : "chain") +
                                           " driver: a neighbour exchange "
                                           "timed out");
            }
        }
        valid = b.nr;
        halted = h;
        if (h >= 0) valid = h + 1;
        if (time_kernels) {
            for (size_t k = 0; k + 1 < b.nev; k += 2) {
                float ms = 0;
                MKH_HIP(hipEventElapsedTime(&ms, ev_pool[k], ev_pool[k + 1]));
                kernel_ms += ms;
            }
        }
        return 0;
    }

    /* Records `valid` rows of the batch as done and sets the outputs of
     * step(): the time reached, and status 0 (more), 1 (done) or 2 (NaN). */
    int close_batch(const Batch& b, long long valid, int halted,
                    double* tout, int* status)
    {
        for (long long k = 0; k < valid; ++k) {
            if (dst_times) dst_times[rows_done + k] = b.tlog[k];
            if (dst_paces) dst_paces[rows_done + k] = b.plog[k];
        }
        rows_done += valid;
        if (halted >= 0) {
            *tout = row_t_after[halted];
            *status = 2;
            running = false;
            return fail(MKH_E_NAN, "Encountered nan in simulation.");
        }
        *tout = plan.t;
        *status = plan.done ? 1 : 0;
        if (plan.done) running = false;
        return 0;
    }

    int step(double* tout, int* status) override
    {
        if (!running) return fail(MKH_E_STATE, "simulation not initialised");
        Batch b;
        int e = plan_batch(b);
        if (e) return e;
        const long long ns = b.ns;
        if (ns == 0) {
            *tout = plan.t;
            *status = 1;
            return 0;
        }
        /* Launch */
        const unsigned bt = (unsigned)cfg.block_threads;
        if (driver == MKH_DRIVER_GRID) {
            for (long long s = 0; s < ns; ++s) {
                e = grid_launch(b, s, nullptr);
                if (e) return e;
            }
        } else if (driver == MKH_DRIVER_SHEET) {
            MkSheetT<R> sh = sheet_args();
            const unsigned ntiles = (unsigned)(sh_tx * sh_ty);
            long long s0 = 0;
            for (long long s = 0; s < ns; ++s) {
                if (h_steps[s].row >= 0 || s == ns - 1) {
                    /* reads state[cur], writes state[cur ^ 1] */
                    Args a = base_args();
                    int i0 = (int)s0, i1 = (int)(s + 1);
                    sh.g0 = (int)sh_steps;
                    void* params[] = {&a, &sh, &i0, &i1};
                    if (time_kernels)
                        MKH_HIP(hipEventRecord(event(b.nev), stream));
                    MKH_HIP(hipModuleLaunchCooperativeKernel(
                        prog->f_sheet, ntiles, 1, 1, 256, 1, 1, 0, stream,
                        params));
                    if (time_kernels) {
                        MKH_HIP(hipEventRecord(event(b.nev + 1), stream));
                        b.nev += 2;
                    }
                    launches++;
                    sh_steps += i1 - i0;
                    cur ^= 1;
                    s0 = s + 1;
                    e = emit_row(b, s);
                    if (e) return e;
                }
            }
        } else if (driver == MKH_DRIVER_CHAIN) {
            MkChainT<R> ch = chain_args();
            long long s0 = 0;
            for (long long s = 0; s < ns; ++s) {
                if (h_steps[s].row >= 0 || s == ns - 1) {
                    /* reads state[cur], writes state[cur ^ 1] */
                    Args a = base_args();
                    int i0 = (int)s0, i1 = (int)(s + 1);
                    ch.g0 = (int)ch_steps;
                    void* params[] = {&a, &ch, &i0, &i1};
                    if (time_kernels)
                        MKH_HIP(hipEventRecord(event(b.nev), stream));
                    MKH_HIP(hipModuleLaunchCooperativeKernel(
                        prog->f_chain, (unsigned)ch_tiles, 1, 1, 64, 1, 1, 0,
                        stream, params));
                    if (time_kernels) {
                        MKH_HIP(hipEventRecord(event(b.nev + 1), stream));
                        b.nev += 2;
                    }
                    launches++;
                    ch_steps += i1 - i0;
                    cur ^= 1;
                    s0 = s + 1;
                    e = emit_row(b, s);
                    if (e) return e;
                }
            }
        } else {
            hipFunction_t f = driver == MKH_DRIVER_CELLS ? prog->f_cells
                                                         : prog->f_block;
            unsigned grid = driver == MKH_DRIVER_CELLS
                                ? (unsigned)((n + bt - 1) / bt)
                                : (unsigned)ndom;
            Args a = base_args();
            R* no_traj = nullptr;
            long long s0 = 0;
            for (long long s = 0; s < ns; ++s) {
                if (h_steps[s].row >= 0 || s == ns - 1) {
                    int i0 = (int)s0, i1 = (int)(s + 1);
                    void* params[] = {&a, &i0, &i1, &no_traj};
                    e = launch(f, grid, params, b.nev);
                    if (e) return e;
                    s0 = s + 1;
                    e = emit_row(b, s);
                    if (e) return e;
                }
            }
        }
        long long valid = 0;
        int halted = -1;
        e = finish_batch(b, valid, halted);
        if (e) return e;
        return close_batch(b, valid, halted, tout, status);
    }

    int clean() override
    {
        running = false;
        return 0;
    }

    int set_option(const std::string& k, long long v) override
    {
        if (k == "batch_steps") {
            if (v < 1) return fail(MKH_E_ARG, "batch_steps must be >= 1");
            batch_steps = v;
        } else if (k == "log_rows") {
            log_rows_cap = v;
        } else if (k == "time_kernels") {
            time_kernels = v != 0;
        } else if (k == "driver") {
            if (cfg.coupling == MKH_COUPLING_GRAPH && !have_graph &&
                (v == MKH_DRIVER_BLOCK || v == MKH_DRIVER_CHAIN)) {
                /* checked against the graph when it arrives */
                forced_driver = (int)v;
                return 0;
            }
            if (v < -1 || v > MKH_DRIVER_CHAIN ||
                (v >= 0 && v != MKH_DRIVER_CHAIN && !feasible((int)v)) ||
                (v == MKH_DRIVER_CHAIN && !is_chain) ||
                (have_graph && v >= 0 && v != MKH_DRIVER_GRID &&
                 v != MKH_DRIVER_BLOCK && v != MKH_DRIVER_CHAIN))
                return fail(MKH_E_ARG, "driver " + std::to_string(v) +
                                           " cannot run this geometry and "
                                           "coupling");
            forced_driver = (int)v;
            if (have_graph) {
                std::vector<long long> i(ci), jj(cj);
                std::vector<double> g(cg);
                return set_connections(i.data(), jj.data(), g.data(),
                                       (long long)i.size());
            }
            choose_driver();
        } else if (k == "graph_copies") {
            if (v < 1) return fail(MKH_E_ARG, "graph_copies must be >= 1");
            graph_copies = v;
            if (have_graph) {
                std::vector<long long> i(ci), jj(cj);
                std::vector<double> g(cg);
                return set_connections(i.data(), jj.data(), g.data(),
                                       (long long)i.size());
            }
        } else if (k == "adj_release") {
            if (v) adj.release();
        } else if (k == "adj_sheet") {
            no_sheet_rev = v == 0;
        } else if (k == "adj_overlap") {
            no_adj_overlap = v == 0;
        } else {
            return fail(MKH_E_ARG, "unknown option " + k);
        }
        return 0;
    }

    int timing(double* ms, long long* nl, long long* cs) override
    {
        if (ms) *ms = kernel_ms;
        if (nl) *nl = launches;
        if (cs) *cs = cell_steps;
        return 0;
    }

    int reset_timing() override
    {
        kernel_ms = 0;
        launches = 0;
        cell_steps = 0;
        return 0;
    }


    unsigned grid_launch_size()
    {
        const unsigned bt = (unsigned)cfg.block_threads;
        if ((cfg.coupling == MKH_COUPLING_STENCIL ||
             cfg.coupling == MKH_COUPLING_HETERO) && cfg.ny > 1) {
            long long tx = (cfg.nx + 63) / 64;
            long long ty = (cfg.ny + (bt / 64) - 1) / (bt / 64);
            return (unsigned)(tx * ty);
        }
        return (unsigned)((n + bt - 1) / bt);
    }

    /* Adjoint work buffers, kept between calls (sizes permitting) and
     * released by the "adj_release" option or with the simulation. */
    /* "adj_sheet" option: 0 keeps the sheet's reverse sweep on the
     * one-step kernel (mk_adj_step) */
    bool no_sheet_rev = false;
    /* "adj_overlap" option: 0 keeps the recomputation of earlier segments
     * in line with the reverse steps on small domains */
    bool no_adj_overlap = false;
    /* second stream (default priority) for overlapped recomputation */
    hipStream_t rc_stream = nullptr;
    hipEvent_t rc_ev[4] = {};

    struct AdjBufs {
        DevBuf<Step> steps;
        DevBuf<R> ckpt, traj, traj2, work, work2, mu[2], ab[2];
        DevBuf<double> gfield, gx, gy, loss, target, trace;
        DevBuf<int> probe_of;
        DevBuf<long long> probes;
        /* the probe list probe_of was built for */
        std::vector<long long> probe_key;
        void release()
        {
            steps.release();
            ckpt.release();
            traj.release();
            traj2.release();
            work.release();
            work2.release();
            for (int k = 0; k < 2; ++k) {
                mu[k].release();
                ab[k].release();
            }
            gfield.release();
            gx.release();
            gy.release();
            loss.release();
            target.release();
            trace.release();
            probe_of.release();
            probes.release();
            probe_key.clear();
        }
    } adj;

    /* Wall time of the last adj_run's phases (ms): forward with
     * checkpoints, recomputation of earlier segments, reverse steps. */
    double adj_ms[3] = {0, 0, 0};
    /* reverse kernel of the last adj_run (MKH_ADJ_REV_*) and its launches */
    int adj_rev_kind = -1;
    long long adj_rev_launches = 0;

    /* Discrete adjoint of a run from the current state (which it leaves
     * unchanged). The forward pass keeps a checkpoint at the start of every
     * segment of K steps and writes the last segment's states straight into
     * the trajectory buffer. The reverse pass then walks the segments from
     * last to first: recompute the segment's states from its checkpoint
     * (not needed for the last), then one fused reverse step per time step.
     * With K >= N nothing is recomputed. */
    int adj_run(double tmin, double tmax, double dt, const double* events,
                long long nev, const long long* probes, long long nprobe,
                const double* target, long long nrows, int K, double* loss,
                double* gfield_out, double* gg_out, double* trace_out) override
    {
        /* trace_out != NULL: forward only, V at the probes after each step
         * into trace_out[s * nprobe + p] (nrows = its row capacity) */
        if (trace_out)
            return trace_run(tmin, tmax, dt, events, nev, probes, nprobe,
                             trace_out, nrows, nullptr, 0);
        /* the chain layout's adjoint runs on the grid driver (its CSR is
         * built with it) */
        DriverAs as_grid(driver, driver == MKH_DRIVER_CHAIN ? MKH_DRIVER_GRID
                                                            : driver);
        if (!prog->f_adj_step)
            return fail(MKH_E_STATE, "kernel was generated without the "
                                     "adjoint");
        if (cfg.coupling == MKH_COUPLING_GRAPH && driver != MKH_DRIVER_GRID &&
            driver != MKH_DRIVER_BLOCK)
            return fail(MKH_E_STATE, "graph adjoint needs the grid or block "
                                     "layout");
        if (K < 1) return fail(MKH_E_ARG, "checkpoint interval must be >= 1");
        /* the reverse kernels index cells with 32-bit integers */
        if (n > INT_MAX)
            return fail(MKH_E_ARG, "the adjoint supports at most 2^31 - 1 "
                                   "cells");
        Planner pl;
        int f = pl.init(tmin, tmax, dt, events, nev, HUGE_VAL);
        if (f) return pacing_error(f);
        std::vector<Step> all;
        StepRec r;
        while (!pl.done) {
            f = pl.next(r);
            if (f) return pacing_error(f);
            Step st;
            st.t = (R)r.t;
            st.dt = (R)r.dt;
            st.pace = (R)r.pace;
            st.row = -1;
            all.push_back(st);
        }
        const long long N = (long long)all.size();
        if (N == 0) return fail(MKH_E_ARG, "empty run");
        if (N > INT_MAX) return fail(MKH_E_ARG, "run too long");
        if (nrows < N)
            return fail(MKH_E_ARG, "target needs one row per time step ("
                                       + std::to_string(N) + ")");
        for (long long k = 0; k < nprobe; ++k)
            if (probes[k] < 0 || probes[k] >= n)
                return fail(MKH_E_ARG, "probe cell out of range");
        {
            /* the loss keeps one probe index per cell (probe_of): a cell
             * listed twice would lose the earlier probe's terms */
            std::vector<long long> ps(probes, probes + nprobe);
            std::sort(ps.begin(), ps.end());
            if (std::adjacent_find(ps.begin(), ps.end()) != ps.end())
                return fail(MKH_E_ARG, "probe cell listed twice");
        }
        const size_t SN = (size_t)cfg.n_state * npad;
        const long long Kl = std::min<long long>(K, N);
        const long long nseg = (N + Kl - 1) / Kl;
        const unsigned bt = (unsigned)cfg.block_threads;
        const unsigned ggrid = grid_launch_size();
        const unsigned cgrid = (unsigned)((n + bt - 1) / bt);
        /* conductance-gradient partial sums: one per workgroup of the
         * reverse kernel (per domain on the block layout) */
        const size_t nparts =
            std::max<size_t>(cgrid, driver == MKH_DRIVER_BLOCK
                                        ? (size_t)std::max<long long>(1, ndom)
                                        : 1);
        DevBuf<Step>& dsteps = adj.steps;
        DevBuf<R>& ckpt = adj.ckpt;
        DevBuf<R>& traj = adj.traj;
        DevBuf<R>& muA = adj.mu[0];
        DevBuf<R>& muB = adj.mu[1];
        DevBuf<R>& abA = adj.ab[0];
        DevBuf<R>& abB = adj.ab[1];
        DevBuf<double>& gfield = adj.gfield;
        DevBuf<double>& gxa = adj.gx;
        DevBuf<double>& gya = adj.gy;
        DevBuf<double>& lossa = adj.loss;
        DevBuf<double>& dtarget = adj.target;
        DevBuf<int>& probe_of = adj.probe_of;
        DevBuf<long long>& dprobes = adj.probes;
        MKH_HIP(dsteps.alloc(N));
        MKH_HIP(hipMemcpy(dsteps.p, all.data(), N * sizeof(Step),
                          hipMemcpyHostToDevice));
        /* trajectory: the longest segment plus its initial state */
        const long long tslots = Kl + 1;
        if (nseg > 1) MKH_HIP(ckpt.alloc((nseg - 1) * SN));
        MKH_HIP(traj.alloc((size_t)tslots * SN));
        MKH_HIP(dprobes.alloc(std::max<long long>(1, nprobe)));
        if (nprobe > 0)
            MKH_HIP(hipMemcpy(dprobes.p, probes, nprobe * sizeof(long long),
                              hipMemcpyHostToDevice));
        MkAdjT<R> j;
        std::memset(&j, 0, sizeof(j));
        j.probes = dprobes.p;
        j.n_probe = (int)nprobe;
        j.target_row = -1;
        {
            MKH_HIP(muA.alloc(SN));
            MKH_HIP(muB.alloc(SN));
            MKH_HIP(abA.alloc(npad));
            MKH_HIP(abB.alloc(npad));
            MKH_HIP(gfield.alloc(
                std::max<size_t>(1, (size_t)cfg.n_field * npad)));
            MKH_HIP(gxa.alloc(nparts));
            MKH_HIP(gya.alloc(nparts));
            MKH_HIP(lossa.alloc(std::max<long long>(1, nprobe)));
            MKH_HIP(hipMemset(gfield.p, 0, gfield.n * sizeof(double)));
            MKH_HIP(hipMemset(gxa.p, 0, nparts * sizeof(double)));
            MKH_HIP(hipMemset(gya.p, 0, nparts * sizeof(double)));
            MKH_HIP(hipMemset(lossa.p, 0, lossa.n * sizeof(double)));
            /* the cell -> probe map (npad ints: 33 MB for 8.4 M cells) is
             * built and uploaded again only when the probes change */
            if (!(probe_of.p && probe_of.n == (size_t)npad &&
                  adj.probe_key.size() == (size_t)nprobe &&
                  std::equal(probes, probes + nprobe,
                             adj.probe_key.begin()))) {
                adj.probe_key.clear();
                std::vector<int> pof(npad, -1);
                for (long long k = 0; k < nprobe; ++k)
                    pof[probes[k]] = (int)k;
                MKH_HIP(probe_of.alloc(npad));
                MKH_HIP(hipMemcpy(probe_of.p, pof.data(), npad * sizeof(int),
                                  hipMemcpyHostToDevice));
                adj.probe_key.assign(probes, probes + nprobe);
            }
            MKH_HIP(dtarget.alloc(std::max<long long>(1, N * nprobe)));
            if (nprobe > 0)
                MKH_HIP(hipMemcpy(dtarget.p, target,
                                  N * nprobe * sizeof(double),
                                  hipMemcpyHostToDevice));
            j.gfield = gfield.p;
            j.gx_acc = gxa.p;
            j.gy_acc = gya.p;
            j.loss_acc = lossa.p;
            j.probe_of = probe_of.p;
            j.target = dtarget.p;
        }
        int m1 = -1;
        MKH_HIP(hipMemcpy(halt.p, &m1, sizeof(int), hipMemcpyHostToDevice));

        Args a = base_args();
        a.steps = dsteps.p;
        auto grid_step = [&](R* src, R* dst, long long si,
                             hipStream_t st) -> int {
            int sii = (int)si;
            void* params[] = {&a, &src, &dst, &sii};
            MKH_HIP(hipModuleLaunchKernel(prog->f_grid, ggrid, 1, 1, bt, 1,
                                          1, 0, st, params, nullptr));
            return 0;
        };
        /* tbase: the trajectory buffer of the segment being reversed */
        R* tbase = traj.p;
        auto slot = [&](long long k) { return tbase + (size_t)k * SN; };
        /* The persistent drivers (sheet, block, cells) advance a scratch
         * copy of the state (they read and write state_a) and store each
         * step's state into the trajectory when asked. The block layout
         * also has a persistent reverse kernel. */
        const bool use_pers = driver == MKH_DRIVER_SHEET ||
                              driver == MKH_DRIVER_BLOCK ||
                              driver == MKH_DRIVER_CELLS;
        const bool block_rev = driver == MKH_DRIVER_BLOCK;
        /* The persistent reverse sheet: 64 x 4 adj_cpt tiles (MK_ACPT in
         * the kernel source), all co-resident; they are no more than the
         * forward sheet's (4 * sheet_cpt rows), whose edge buffer they use */
        const int rtx = sh_tx;
        const int ath = 4 * std::max(cfg.adj_cpt, 1);
        const int rty = (int)((cfg.ny + ath - 1) / ath);
        bool sheet_rev = false;
        if (driver == MKH_DRIVER_SHEET && cfg.adj_cpt > 0 &&
            prog->f_adj_sheet &&
            sh_edges.p && (size_t)rtx * rty * 4 * 64 * 2 * XW <= sh_edges.n &&
            (size_t)rtx * rty <= nparts) {
            int per_cu = 0;
            MKH_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, prog->f_adj_sheet, 256, 0));
            hipDeviceProp_t prop;
            MKH_HIP(hipGetDeviceProperties(&prop, cfg.device));
            sheet_rev = (long long)per_cu * prop.multiProcessorCount >=
                        (long long)rtx * rty && !no_sheet_rev;
        }
        /* Small domains on the one-step kernels: segment sg - 1 is
         * recomputed on a second stream while the reverse steps of segment
         * sg run, the two segments' states in two trajectory buffers. Only
         * where the workgroups of both kernels fit on the device at once
         * (otherwise each would only slow the other). Same kernels, same
         * inputs: the results are those of the in-line order. */
        int ncu = 0;
        MKH_HIP(hipDeviceGetAttribute(
            &ncu, hipDeviceAttributeMultiprocessorCount, cfg.device));
        const bool overlap = !use_pers && nseg > 1 && !no_adj_overlap &&
                             (long long)ggrid + cgrid <= ncu;
        if (overlap) {
            MKH_HIP(adj.traj2.alloc((size_t)tslots * SN));
            if (!rc_stream) {
                MKH_HIP(hipStreamCreate(&rc_stream));
                for (auto& e : rc_ev) MKH_HIP(hipEventCreate(&e));
            }
        }
        adj_rev_kind = sheet_rev ? MKH_ADJ_REV_SHEET
                       : block_rev ? MKH_ADJ_REV_BLOCK
                       : overlap ? MKH_ADJ_REV_STEP_OVERLAP
                                 : MKH_ADJ_REV_STEP;
        adj_rev_launches = 0;
        DevBuf<R>& work = adj.work;
        DevBuf<R>& work2 = adj.work2;
        if (use_pers) MKH_HIP(work.alloc(SN));
        if (use_pers && driver == MKH_DRIVER_SHEET)
            MKH_HIP(work2.alloc(SN));
        /* wcur: the scratch state the persistent drivers advance (the sheet
         * writes its result to the other buffer) */
        R* wcur = work.p;
        auto pers_run = [&](long long s0, long long s1, R* trajp) -> int {
            Args as = a;
            as.state_a = wcur;
            int i0 = (int)s0, i1 = (int)s1;
            if (driver == MKH_DRIVER_SHEET) {
                R* wnext = wcur == work.p ? work2.p : work.p;
                as.state_b = wnext;
                MkSheetT<R> sh = sheet_args();
                sh.g0 = (int)sh_steps;
                sh.traj = trajp;
                void* params[] = {&as, &sh, &i0, &i1};
                MKH_HIP(hipModuleLaunchCooperativeKernel(
                    prog->f_sheet, (unsigned)(sh_tx * sh_ty), 1, 1, 256, 1,
                    1, 0, stream, params));
                sh_steps += s1 - s0;
                wcur = wnext;
            } else {
                void* params[] = {&as, &i0, &i1, &trajp};
                const bool blk = driver == MKH_DRIVER_BLOCK;
                MKH_HIP(hipModuleLaunchKernel(
                    blk ? prog->f_block : prog->f_cells,
                    blk ? (unsigned)ndom : cgrid, 1, 1, bt, 1, 1, 0, stream,
                    params, nullptr));
            }
            return 0;
        };
        hipEvent_t ev[4];
        for (int k = 0; k < 4; ++k) MKH_HIP(hipEventCreate(&ev[k]));
        struct EvFree {
            hipEvent_t* e;
            ~EvFree()
            {
                for (int k = 0; k < 4; ++k) (void)hipEventDestroy(e[k]);
            }
        } evfree{ev};
        /* whatever way this call ends (an error return included), nothing
         * it queued on the second stream outlives it (a later call reuses
         * the trajectory buffers) or the events above (destroyed after
         * this, being declared before it) */
        struct RcDrain {
            hipStream_t st;
            ~RcDrain()
            {
                if (st) (void)hipStreamSynchronize(st);
            }
        } rc_drain{overlap ? rc_stream : nullptr};
        MKH_HIP(hipEventRecord(ev[0], stream));

        /* Forward */
        R* wa = slot(0);
        R* wb = slot(1);
        MKH_HIP(hipMemcpyAsync(wa, state[cur].p, SN * sizeof(R),
                               hipMemcpyDeviceToDevice, stream));
        if (use_pers)
            MKH_HIP(hipMemcpyAsync(wcur, state[cur].p, SN * sizeof(R),
                                   hipMemcpyDeviceToDevice, stream));
        for (long long sg = 0; use_pers && sg < nseg; ++sg) {
            const long long s0 = sg * Kl, s1 = std::min(N, s0 + Kl);
            R* keep = sg < nseg - 1 ? ckpt.p + sg * SN : slot(0);
            MKH_HIP(hipMemcpyAsync(keep, wcur, SN * sizeof(R),
                                   hipMemcpyDeviceToDevice, stream));
            int e = pers_run(s0, s1, sg < nseg - 1 ? nullptr : slot(1));
            if (e) return e;
        }
        for (long long sg = 0; !use_pers && sg < nseg; ++sg) {
            const long long s0 = sg * Kl, s1 = std::min(N, s0 + Kl);
            if (sg < nseg - 1) {
                MKH_HIP(hipMemcpyAsync(ckpt.p + sg * SN, wa, SN * sizeof(R),
                                       hipMemcpyDeviceToDevice, stream));
                for (long long si = s0; si < s1; ++si) {
                    int e = grid_step(wa, wb, si, stream);
                    if (e) return e;
                    std::swap(wa, wb);
                }
            } else {
                /* last segment: keep every state */
                if (wa != slot(0))
                    MKH_HIP(hipMemcpyAsync(slot(0), wa, SN * sizeof(R),
                                           hipMemcpyDeviceToDevice, stream));
                for (long long si = s0; si < s1; ++si) {
                    int e = grid_step(slot(si - s0), slot(si - s0 + 1), si,
                                      stream);
                    if (e) return e;
                }
            }
        }
        MKH_HIP(hipEventRecord(ev[1], stream));

        /* Reverse */
        R* mu_in = nullptr;
        R* ab_in = nullptr;
        R* mu_out = muA.p;
        R* ab_out = abA.p;
        auto reverse_step = [&](const R* yn, const R* yn1, long long si,
                                int row) -> int {
            MkAdjT<R> js = j;
            js.yn = yn;
            js.yn1 = yn1;
            js.mu_in = mu_in;
            js.ab_in = ab_in;
            js.mu_out = mu_out;
            js.ab_out = ab_out;
            js.target_row = row;
            int sii = (int)si;
            void* params[] = {&a, &js, &sii};
            MKH_HIP(hipModuleLaunchKernel(prog->f_adj_step, cgrid, 1, 1, bt,
                                          1, 1, 0, stream, params, nullptr));
            adj_rev_launches++;
            mu_in = mu_out;
            ab_in = ab_out;
            mu_out = (mu_out == muA.p) ? muB.p : muA.p;
            ab_out = (ab_out == abA.p) ? abB.p : abA.p;
            return 0;
        };
        double recompute_ms = 0;
        /* overlapped order: segment sg's states are in bufs[(nseg - 1 - sg)
         * % 2]; the recomputation of sg - 1 (second stream) waits for the
         * reverse steps of sg + 1 to release its buffer, the reverse steps
         * of sg for the recomputation of sg. rc_ev[2 * parity] and
         * rc_ev[2 * parity + 1] bracket one recomputation, read back
         * (host-side) before they are recorded again. */
        R* bufs[2] = {traj.p, adj.traj2.p};
        hipEvent_t freed = ev[1];       /* forward done: checkpoints ready */
        auto rc_collect = [&](int par) -> int {
            MKH_HIP(hipEventSynchronize(rc_ev[2 * par + 1]));
            float ms = 0;
            MKH_HIP(hipEventElapsedTime(&ms, rc_ev[2 * par],
                                        rc_ev[2 * par + 1]));
            recompute_ms += ms;
            return 0;
        };
        bool rc_pending[2] = {false, false};
        for (long long sg = nseg - 1; overlap && sg >= 0; --sg) {
            const long long s0 = sg * Kl, s1 = std::min(N, s0 + Kl);
            if (sg > 0) {
                /* recompute segment sg - 1 into the other buffer */
                const long long p0 = (sg - 1) * Kl;
                const int par = (int)((nseg - sg) % 2);
                R* base = bufs[par];
                if (rc_pending[par]) {
                    int e = rc_collect(par);
                    if (e) return e;
                }
                MKH_HIP(hipStreamWaitEvent(rc_stream, freed, 0));
                MKH_HIP(hipEventRecord(rc_ev[2 * par], rc_stream));
                MKH_HIP(hipMemcpyAsync(base, ckpt.p + (sg - 1) * SN,
                                       SN * sizeof(R),
                                       hipMemcpyDeviceToDevice, rc_stream));
                for (long long si = p0; si < s0; ++si) {
                    int e = grid_step(base + (size_t)(si - p0) * SN,
                                      base + (size_t)(si - p0 + 1) * SN, si,
                                      rc_stream);
                    if (e) return e;
                }
                MKH_HIP(hipEventRecord(rc_ev[2 * par + 1], rc_stream));
                rc_pending[par] = true;
            }
            /* reverse steps of segment sg, once its states are there */
            const int cpar = (int)((nseg - 1 - sg) % 2);
            tbase = bufs[cpar];
            if (sg < nseg - 1)
                MKH_HIP(hipStreamWaitEvent(stream, rc_ev[2 * cpar + 1], 0));
            for (long long si = s1 - 1; si >= s0; --si) {
                int e = reverse_step(slot(si - s0), slot(si - s0 + 1), si,
                                     (int)si);
                if (e) return e;
            }
            /* its buffer is free for segment sg - 2 after these steps */
            MKH_HIP(hipEventRecord(ev[2], stream));
            freed = ev[2];
        }
        for (int par = 0; par < 2; ++par)
            if (rc_pending[par]) {
                int e = rc_collect(par);
                if (e) return e;
            }
        for (long long sg = nseg - 1; !overlap && sg >= 0; --sg) {
            const long long s0 = sg * Kl, s1 = std::min(N, s0 + Kl);
            if (sg < nseg - 1) {
                MKH_HIP(hipEventRecord(ev[2], stream));
                MKH_HIP(hipMemcpyAsync(slot(0), ckpt.p + sg * SN,
                                       SN * sizeof(R),
                                       hipMemcpyDeviceToDevice, stream));
                if (use_pers) {
                    MKH_HIP(hipMemcpyAsync(wcur, ckpt.p + sg * SN,
                                           SN * sizeof(R),
                                           hipMemcpyDeviceToDevice, stream));
                    int e = pers_run(s0, s1, slot(1));
                    if (e) return e;
                } else {
                    for (long long si = s0; si < s1; ++si) {
                        int e = grid_step(slot(si - s0), slot(si - s0 + 1),
                                          si, stream);
                        if (e) return e;
                    }
                }
                MKH_HIP(hipEventRecord(ev[3], stream));
                MKH_HIP(hipEventSynchronize(ev[3]));
                float ms = 0;
                MKH_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
                recompute_ms += ms;
            }
            if (sheet_rev) {
                /* every reverse step of the segment in one cooperative
                 * launch; mu in LDS, carried in muA and abar in abA */
                MkAdjT<R> js = j;
                js.yn = slot(0);
                js.mu_out = muA.p;
                js.ab_out = abA.p;
                MkSheetT<R> shr = sheet_args();
                shr.tiles_y = rty;
                shr.g0 = 0;
                int i0 = (int)s0, i1 = (int)s1;
                int flags = (sg == nseg - 1 ? 1 : 0) | (sg == 0 ? 2 : 0);
                /* the reverse tags start at 1 in every launch: clear the
                 * edge words (the forward's tags are larger, and they stay
                 * above the reverse's: tags only grow until a run's init
                 * clears the words) */
                MKH_HIP(hipMemsetAsync(sh_edges.p, 0,
                                       sh_edges.n * sizeof(unsigned long long),
                                       stream));
                void* params[] = {&a, &js, &shr, &i0, &i1, &flags};
                MKH_HIP(hipModuleLaunchCooperativeKernel(
                    prog->f_adj_sheet, (unsigned)(rtx * rty), 1, 1, 256, 1,
                    1, 0, stream, params));
                adj_rev_launches++;
                continue;
            }
            if (block_rev) {
                /* every reverse step of the segment in one launch; mu in
                 * place in muA, abar carried in abA */
                MkAdjT<R> js = j;
                js.yn = slot(0);
                js.mu_out = muA.p;
                js.ab_out = abA.p;
                int i0 = (int)s0, i1 = (int)s1;
                int flags = (sg == nseg - 1 ? 1 : 0) | (sg == 0 ? 2 : 0);
                void* params[] = {&a, &js, &i0, &i1, &flags};
                MKH_HIP(hipModuleLaunchKernel(prog->f_adj_block,
                                              (unsigned)ndom, 1, 1, bt, 1, 1,
                                              0, stream, params, nullptr));
                adj_rev_launches++;
                continue;
            }
            for (long long si = s1 - 1; si >= s0; --si) {
                int e = reverse_step(slot(si - s0), slot(si - s0 + 1), si,
                                     (int)si);
                if (e) return e;
            }
        }
        /* closing pass: conductance terms of step 0 (y_0 is slot 0) */
        if (!block_rev && !sheet_rev) {
            int e = reverse_step(nullptr, slot(0), 0, -1);
            if (e) return e;
        }
        MKH_HIP(hipEventRecord(ev[3], stream));
        MKH_HIP(hipEventSynchronize(ev[3]));
        if (use_pers && driver == MKH_DRIVER_SHEET) {
            int e = 0;
            MKH_HIP(hipMemcpy(&e, sh_err.p, sizeof(int),
                              hipMemcpyDeviceToHost));
            if (e)
                return fail(MKH_E_HIP, "sheet driver: a neighbour exchange "
                                       "timed out");
        }
        {
            float f0 = 0, f1 = 0;
            MKH_HIP(hipEventElapsedTime(&f0, ev[0], ev[1]));
            MKH_HIP(hipEventElapsedTime(&f1, ev[1], ev[3]));
            adj_ms[0] = f0;
            adj_ms[1] = recompute_ms;
            /* overlapped: the recomputation ran during the reverse phase */
            adj_ms[2] = overlap ? f1 : f1 - recompute_ms;
        }
        /* Reductions in double precision, in cell / workgroup order */
        if (gfield_out && cfg.n_field > 0) {
            if (cell_order.p && prog->f_adj_gather) {
                /* gathered into the caller's order on the device, then
                 * one copy into the caller's buffer */
                const long long tot = n * cfg.n_field;
                MKH_HIP(gsorted.alloc((size_t)tot));
                const double* src = gfield.p;
                const long long* ord = cell_order.p;
                double* dst = gsorted.p;
                long long nn = n, np = npad;
                int nf = cfg.n_field;
                void* params[] = {&src, &ord, &dst, &nn, &np, &nf};
                const unsigned gg = (unsigned)std::min<long long>(
                    (tot + 255) / 256, 16384);
                MKH_HIP(hipModuleLaunchKernel(prog->f_adj_gather, gg, 1, 1,
                                              256, 1, 1, 0, stream, params,
                                              nullptr));
                MKH_HIP(hipMemcpyAsync(gfield_out, dst,
                                       (size_t)tot * sizeof(double),
                                       hipMemcpyDeviceToHost, stream));
                MKH_HIP(hipStreamSynchronize(stream));
            } else if (npad == n) {
                /* same layout: straight into the caller's buffer */
                MKH_HIP(hipMemcpy(gfield_out, gfield.p,
                                  gfield.n * sizeof(double),
                                  hipMemcpyDeviceToHost));
            } else {
                std::vector<double> hf(gfield.n);
                MKH_HIP(hipMemcpy(hf.data(), gfield.p,
                                  gfield.n * sizeof(double),
                                  hipMemcpyDeviceToHost));
                for (int k = 0; k < cfg.n_field; ++k)
                    for (long long c = 0; c < n; ++c)
                        gfield_out[k * n + c] = hf[(size_t)k * npad + c];
            }
        }
        /* the conductance partial sums (one per workgroup or domain) */
        std::vector<double> h(nparts);
        double sx = 0, sy = 0, sl = 0;
        MKH_HIP(hipMemcpy(h.data(), gxa.p, nparts * sizeof(double),
                          hipMemcpyDeviceToHost));
        for (size_t b = 0; b < nparts; ++b) sx += h[b];
        MKH_HIP(hipMemcpy(h.data(), gya.p, nparts * sizeof(double),
                          hipMemcpyDeviceToHost));
        for (size_t b = 0; b < nparts; ++b) sy += h[b];
        std::vector<double> hl(lossa.n);
        MKH_HIP(hipMemcpy(hl.data(), lossa.p, lossa.n * sizeof(double),
                          hipMemcpyDeviceToHost));
        for (long long k = 0; k < nprobe; ++k) sl += hl[k];
        if (gg_out) {
            gg_out[0] = sx;
            gg_out[1] = sy;
        }
        if (loss) *loss = sl;
        return 0;
    }

    /* Forward-only run of mkh_adj_trace and mkh_adj_trace_static: V at the
     * probes after each of the N time steps of tmin..tmax, into trace_out
     * (host, trace_out[s * nprobe + p]). The run advances a copy of the
     * state, which it leaves unchanged. The device buffers (step table,
     * state copies, probe list, trace) are the adjoint's work buffers, or,
     * when `arena` is given, carved from that block of `arena_bytes` bytes:
     * then nothing is allocated, and a run that needs more fails with
     * MKH_E_ARG before anything is launched. */
    int trace_run(double tmin, double tmax, double dt, const double* events,
                  long long nev, const long long* probes, long long nprobe,
                  double* trace_out, long long nrows, char* arena,
                  size_t arena_bytes)
    {
        if (!prog->f_adj_probe)
            return fail(MKH_E_STATE, "kernel was generated without the "
                                     "adjoint");
        DriverAs as_grid(driver, driver == MKH_DRIVER_CHAIN ? MKH_DRIVER_GRID
                                                            : driver);
        if (cfg.coupling == MKH_COUPLING_GRAPH && driver != MKH_DRIVER_GRID &&
            driver != MKH_DRIVER_BLOCK)
            return fail(MKH_E_STATE, "graph adjoint needs the grid or block "
                                     "layout");
        if (n > INT_MAX)
            return fail(MKH_E_ARG, "the adjoint supports at most 2^31 - 1 "
                                   "cells");
        /* count the steps (the planner is deterministic) */
        Planner pl;
        int f = pl.init(tmin, tmax, dt, events, nev, HUGE_VAL);
        if (f) return pacing_error(f);
        long long N = 0;
        {
            Planner dry = pl;
            StepRec r;
            while (!dry.done) {
                f = dry.next(r);
                if (f) return pacing_error(f);
                N++;
            }
        }
        if (N == 0) return fail(MKH_E_ARG, "empty run");
        if (N > INT_MAX) return fail(MKH_E_ARG, "run too long");
        if (nrows < N)
            return fail(MKH_E_ARG, "trace needs one row per time step ("
                                       + std::to_string(N) + ")");
        for (long long k = 0; k < nprobe; ++k)
            if (probes[k] < 0 || probes[k] >= n)
                return fail(MKH_E_ARG, "probe cell out of range");
        const size_t SN = (size_t)cfg.n_state * npad;
        const unsigned bt = (unsigned)cfg.block_threads;
        const unsigned ggrid = grid_launch_size();
        const unsigned cgrid = (unsigned)((n + bt - 1) / bt);
        /* The persistent drivers (sheet, block, cells) advance one scratch
         * state in place (the sheet: two, it writes its result to the
         * other); the grid driver steps between two buffers. */
        const bool use_pers = driver == MKH_DRIVER_SHEET ||
                              driver == MKH_DRIVER_BLOCK ||
                              driver == MKH_DRIVER_CELLS;
        const int nstate = use_pers && driver != MKH_DRIVER_SHEET ? 1 : 2;
        const size_t ntr = (size_t)std::max<long long>(1, N * nprobe);
        const size_t npr = (size_t)std::max<long long>(1, nprobe);
        Step* dsteps = nullptr;
        R* st[2] = {nullptr, nullptr};
        long long* dprobes = nullptr;
        double* dtrace = nullptr;
        if (arena) {
            /* regions on 256-byte boundaries, in this order */
            size_t off = 0;
            auto carve = [&](size_t bytes) {
                char* p = arena + off;
                off += (bytes + 255) / 256 * 256;
                return p;
            };
            dsteps = (Step*)carve((size_t)N * sizeof(Step));
            for (int k = 0; k < nstate; ++k) st[k] = (R*)carve(SN * sizeof(R));
            dprobes = (long long*)carve(npr * sizeof(long long));
            dtrace = (double*)carve(ntr * sizeof(double));
            if (off > arena_bytes)
                return fail(MKH_E_ARG,
                            "trace needs " + std::to_string(off) +
                                " bytes of scratch; the static scratch holds "
                                + std::to_string(arena_bytes));
        } else {
            MKH_HIP(adj.steps.alloc((size_t)N));
            MKH_HIP(adj.traj.alloc((size_t)nstate * SN));
            MKH_HIP(adj.probes.alloc(npr));
            MKH_HIP(adj.trace.alloc(ntr));
            dsteps = adj.steps.p;
            st[0] = adj.traj.p;
            st[1] = nstate > 1 ? adj.traj.p + SN : nullptr;
            dprobes = adj.probes.p;
            dtrace = adj.trace.p;
        }
        /* step table, uploaded in chunks from the stack */
        {
            Step chunk[512];
            long long k = 0, base = 0;
            StepRec r;
            while (!pl.done) {
                f = pl.next(r);
                if (f) return pacing_error(f);
                Step& s1 = chunk[k++];
                s1.t = (R)r.t;
                s1.dt = (R)r.dt;
                s1.pace = (R)r.pace;
                s1.row = -1;
                if (k == 512 || pl.done) {
                    MKH_HIP(hipMemcpy(dsteps + base, chunk,
                                      (size_t)k * sizeof(Step),
                                      hipMemcpyHostToDevice));
                    base += k;
                    k = 0;
                }
            }
        }
        if (nprobe > 0)
            MKH_HIP(hipMemcpy(dprobes, probes,
                              (size_t)nprobe * sizeof(long long),
                              hipMemcpyHostToDevice));
        const int m1 = -1;
        MKH_HIP(hipMemcpy(halt.p, &m1, sizeof(int), hipMemcpyHostToDevice));
        MkAdjT<R> j;
        std::memset(&j, 0, sizeof(j));
        j.probes = dprobes;
        j.n_probe = (int)nprobe;
        j.target_row = -1;
        j.trace = dtrace;
        Args a = base_args();
        a.steps = dsteps;
        R* wa = st[0];
        R* wb = st[1];
        MKH_HIP(hipMemcpyAsync(wa, state[cur].p, SN * sizeof(R),
                               hipMemcpyDeviceToDevice, stream));
        const unsigned pgrid = (unsigned)((nprobe + 63) / 64);
        for (long long si = 0; si < N; ++si) {
            int sii = (int)si;
            if (driver == MKH_DRIVER_SHEET) {
                Args as = a;
                as.state_a = wa;
                as.state_b = wb;
                MkSheetT<R> sh = sheet_args();
                sh.g0 = (int)sh_steps;
                int i0 = sii, i1 = sii + 1;
                void* params[] = {&as, &sh, &i0, &i1};
                MKH_HIP(hipModuleLaunchCooperativeKernel(
                    prog->f_sheet, (unsigned)(sh_tx * sh_ty), 1, 1, 256, 1,
                    1, 0, stream, params));
                sh_steps += 1;
                std::swap(wa, wb);
            } else if (use_pers) {
                Args as = a;
                as.state_a = wa;
                int i0 = sii, i1 = sii + 1;
                R* trajp = nullptr;
                void* params[] = {&as, &i0, &i1, &trajp};
                const bool blk = driver == MKH_DRIVER_BLOCK;
                MKH_HIP(hipModuleLaunchKernel(
                    blk ? prog->f_block : prog->f_cells,
                    blk ? (unsigned)ndom : cgrid, 1, 1, bt, 1, 1, 0, stream,
                    params, nullptr));
            } else {
                void* params[] = {&a, &wa, &wb, &sii};
                MKH_HIP(hipModuleLaunchKernel(prog->f_grid, ggrid, 1, 1, bt,
                                              1, 1, 0, stream, params,
                                              nullptr));
                std::swap(wa, wb);
            }
            if (nprobe > 0) {
                MkAdjT<R> jt = j;
                jt.yn = wa;
                void* params[] = {&a, &jt, &sii};
                MKH_HIP(hipModuleLaunchKernel(prog->f_adj_probe, pgrid, 1, 1,
                                              64, 1, 1, 0, stream, params,
                                              nullptr));
            }
        }
        MKH_HIP(hipStreamSynchronize(stream));
        if (driver == MKH_DRIVER_SHEET) {
            int e = 0;
            MKH_HIP(hipMemcpy(&e, sh_err.p, sizeof(int),
                              hipMemcpyDeviceToHost));
            if (e)
                return fail(MKH_E_HIP, "sheet driver: a neighbour exchange "
                                       "timed out");
        }
        if (nprobe > 0)
            MKH_HIP(hipMemcpy(trace_out, dtrace,
                              (size_t)N * nprobe * sizeof(double),
                              hipMemcpyDeviceToHost));
        return 0;
    }

    /* mkh_adj_trace_static: trace_run on this device's static scratch */
    int trace_static(double tmin, double tmax, double dt,
                     const double* events, long long nev,
                     const long long* probes, long long nprobe,
                     double* trace_out, long long nrows) override
    {
        void* base = nullptr;
        int e = scratch_address(cfg.device, &base);
        if (e) return e;
        return trace_run(tmin, tmax, dt, events, nev, probes, nprobe,
                         trace_out, nrows, (char*)base,
                         (size_t)MKH_STATIC_SCRATCH_BYTES);
    }

    int adj_timing(double* ms) override
    {
        for (int k = 0; k < 3; ++k) ms[k] = adj_ms[k];
        return 0;
    }

    int adj_reverse(int* kind, long long* launches) override
    {
        *kind = adj_rev_kind;
        *launches = adj_rev_launches;
        return 0;
    }

    /* Device cell order[c] of the caller's cell c, for the field gradients
     * of adj_run (empty: the caller's order is the device's). */
    DevBuf<long long> cell_order;
    DevBuf<double> gsorted;

    int set_cell_order(const long long* order, long long m) override
    {
        if (!order || m == 0) {
            cell_order.release();
            gsorted.release();
            return 0;
        }
        if (m != n)
            return fail(MKH_E_ARG, "cell order: one entry per cell needed");
        for (long long c = 0; c < m; ++c)
            if (order[c] < 0 || order[c] >= n)
                return fail(MKH_E_ARG, "cell order: index out of range");
        MKH_HIP(cell_order.alloc((size_t)n));
        MKH_HIP(hipMemcpy(cell_order.p, order, (size_t)n * sizeof(long long),
                          hipMemcpyHostToDevice));
        return 0;
    }

    long long log_rows() override { return rows_done; }
    int get_driver() override { return driver; }
};

}  // namespace

/* -------------------------------------------------------------------------
 * C ABI
 * ---------------------------------------------------------------------- */
/* -------------------------------------------------------------------------
 * Fiber-tissue driver (FiberTissueSimulation's sim_step,
 * _sim/fiber_tissue.c:1003-1160): two simulations, each with its own model
 * and kernel program, advanced in lockstep on one stream with the grid
 * driver. Per step: the junction current from both potentials at t
 * (mk_junction), then the fiber step and the tissue step, each adding its
 * junction current after its stencil. Both simulations were initialised
 * with the same run, so their planners give the same steps.
 * ---------------------------------------------------------------------- */
struct FtBase {
    virtual ~FtBase() {}
    virtual int step(double*, int*) = 0;
};

template <typename R>
struct FtSim : FtBase {
    Sim<R>* f = nullptr;
    Sim<R>* t = nullptr;
    DevBuf<R> jf, jt;
    long long nfx = 0, nfy = 0, ntx = 0, ctx = 0, cty = 0;
    R gft = 0;

    int create(Sim<R>* fs, Sim<R>* ts, long long ctx_, long long cty_,
               double g)
    {
        f = fs;
        t = ts;
        if (f->cfg.coupling != MKH_COUPLING_STENCIL ||
            t->cfg.coupling != MKH_COUPLING_STENCIL)
            return fail(MKH_E_ARG, "fiber and tissue need stencil coupling");
        if (f->cfg.device != t->cfg.device)
            return fail(MKH_E_ARG, "fiber and tissue on different devices");
        nfx = f->cfg.nx;
        nfy = f->cfg.ny;
        ntx = t->cfg.nx;
        ctx = ctx_;
        cty = cty_;
        gft = (R)g;
        if (ctx < 0 || ctx >= ntx || cty < 0 || cty + nfy > t->cfg.ny)
            return fail(MKH_E_ARG, "junction outside the tissue");
        MKH_HIP(hipSetDevice(f->cfg.device));
        MKH_HIP(jf.alloc((size_t)f->npad));
        MKH_HIP(jt.alloc((size_t)t->npad));
        MKH_HIP(hipMemset(jf.p, 0, jf.n * sizeof(R)));
        MKH_HIP(hipMemset(jt.p, 0, jt.n * sizeof(R)));
        return 0;
    }

    int step(double* tout, int* status) override
    {
        if (!f->running || !t->running)
            return fail(MKH_E_STATE, "simulation not initialised");
        if (f->driver != MKH_DRIVER_GRID || t->driver != MKH_DRIVER_GRID)
            return fail(MKH_E_STATE, "fiber-tissue steps need the grid "
                                     "driver");
        MKH_HIP(hipSetDevice(f->cfg.device));
        /* the tissue works on the fiber's stream for this batch */
        hipStream_t own = t->stream;
        MKH_HIP(hipStreamSynchronize(own));
        MKH_HIP(hipStreamSynchronize(f->stream));
        t->stream = f->stream;
        struct Restore {
            Sim<R>* s;
            hipStream_t st;
            ~Restore() { s->stream = st; }
        } restore{t, own};
        const long long cap = std::min(f->log_rows_cap_run,
                                       t->log_rows_cap_run);
        f->log_rows_cap_run = t->log_rows_cap_run = cap;
        f->batch_steps = t->batch_steps = std::min(f->batch_steps,
                                                   t->batch_steps);
        typename Sim<R>::Batch bf, bt;
        int e = f->plan_batch(bf);
        if (e) return e;
        e = t->plan_batch(bt);
        if (e) return e;
        if (bf.ns != bt.ns || bf.nr != bt.nr)
            return fail(MKH_E_STATE, "fiber and tissue runs differ");
        if (bf.ns == 0) {
            *tout = f->plan.t;
            *status = 1;
            return 0;
        }
        const unsigned jgrid = (unsigned)((nfy + 63) / 64);
        for (long long s = 0; s < bf.ns; ++s) {
            const R* vf = f->state[f->cur].p + (long long)f->cfg.i_vm * f->npad;
            const R* vt = t->state[t->cur].p + (long long)t->cfg.i_vm * t->npad;
            R* pjf = jf.p;
            R* pjt = jt.p;
            void* params[] = {&vf, &vt, &pjf, &pjt, &nfx, &nfy, &ntx, &ctx,
                              &cty, &gft};
            MKH_HIP(hipModuleLaunchKernel(f->prog->f_junction, jgrid, 1, 1,
                                          64, 1, 1, 0, f->stream, params,
                                          nullptr));
            e = f->grid_launch(bf, s, jf.p);
            if (e) return e;
            e = t->grid_launch(bt, s, jt.p);
            if (e) return e;
        }
        long long vf_rows = 0, vt_rows = 0;
        int hf = -1, ht = -1;
        e = f->finish_batch(bf, vf_rows, hf);
        if (e) return e;
        e = t->finish_batch(bt, vt_rows, ht);
        if (e) return e;
        /* a NaN in either part ends both at the same row (halt_sim) */
        int h = hf;
        if (ht >= 0 && (h < 0 || ht < h)) h = ht;
        const long long valid = h >= 0 ? h + 1 : bf.nr;
        double t2 = 0;
        int s2 = 0;
        e = t->close_batch(bt, valid, h, &t2, &s2);
        const int ef = f->close_batch(bf, valid, h, tout, status);
        return ef ? ef : e;
    }
};

extern "C" {

const char* mkh_version(void) { return "myokit_amd 0.1"; }

const char* mkh_last_error(void) { return g_err.c_str(); }

int mkh_device_count(int* n)
{
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) c = 0;
    if (n) *n = c;
    return 0;
}

int mkh_device_info(int device, char* name, int name_len,
                    long long* total_mem, int* n_cu)
{
    hipDeviceProp_t p;
    MKH_HIP(hipGetDeviceProperties(&p, device));
    if (name && name_len > 0) {
        std::snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (total_mem) *total_mem = (long long)p.totalGlobalMem;
    if (n_cu) *n_cu = p.multiProcessorCount;
    return 0;
}

int mkh_program_build(const char* source, const char* options,
                      const char* cache_dir, int device, void** program,
                      char* build_log, int build_log_len)
{
    if (!source || !program) return fail(MKH_E_ARG, "null argument");
    std::string arch = "gfx950";
    if (device >= 0) {
        hipDeviceProp_t p;
        MKH_HIP(hipGetDeviceProperties(&p, device));
        arch = p.gcnArchName;
        size_t colon = arch.find(':');
        if (colon != std::string::npos) arch = arch.substr(0, colon);
    }
    std::vector<std::string> opts = split_options(options);
    opts.push_back("--offload-arch=" + arch);
    std::string key = std::string(source);
    for (auto& o : opts) key += "\x01" + o;
    unsigned long long h = fnv1a(key, 1469598103934665603ULL);
    char hex[32];
    std::snprintf(hex, sizeof(hex), "%016llx", h);
    std::string cpath;
    if (cache_dir && cache_dir[0])
        cpath = std::string(cache_dir) + "/mk_" + hex + "_" + arch + ".co";

    std::unique_ptr<Program> prog(new Program());
    prog->device = device;
    bool have = false;
    if (!cpath.empty()) {
        std::ifstream f(cpath, std::ios::binary);
        if (f) {
            prog->code.assign(std::istreambuf_iterator<char>(f),
                              std::istreambuf_iterator<char>());
            have = !prog->code.empty();
        }
    }
    if (!have) {
        hiprtcProgram rp;
        hiprtcResult r =
            hiprtcCreateProgram(&rp, source, "mk_kernel.hip", 0, nullptr,
                                nullptr);
        if (r != HIPRTC_SUCCESS)
            return fail(MKH_E_COMPILE, hiprtcGetErrorString(r));
        std::vector<const char*> cop;
        for (auto& o : opts) cop.push_back(o.c_str());
        r = hiprtcCompileProgram(rp, (int)cop.size(), cop.data());
        size_t lsz = 0;
        hiprtcGetProgramLogSize(rp, &lsz);
        std::string log(lsz, '\0');
        if (lsz) hiprtcGetProgramLog(rp, &log[0]);
        if (build_log && build_log_len > 0) {
            std::snprintf(build_log, build_log_len, "%s", log.c_str());
        }
        if (r != HIPRTC_SUCCESS) {
            hiprtcDestroyProgram(&rp);
            return fail(MKH_E_COMPILE,
                        std::string("hiprtc compile failed: ") +
                            hiprtcGetErrorString(r) + "\n" + log);
        }
        size_t csz = 0;
        hiprtcGetCodeSize(rp, &csz);
        prog->code.resize(csz);
        hiprtcGetCode(rp, prog->code.data());
        hiprtcDestroyProgram(&rp);
        if (!cpath.empty()) {
            std::string tmp = cpath + ".tmp" + std::to_string((long long)getpid());
            std::ofstream f(tmp, std::ios::binary);
            if (f) {
                f.write(prog->code.data(), (std::streamsize)prog->code.size());
                f.close();
                std::rename(tmp.c_str(), cpath.c_str());
            }
        }
    }
    if (device >= 0) {
        MKH_HIP(hipSetDevice(device));
        MKH_HIP(hipModuleLoadData(&prog->module, prog->code.data()));
        MKH_HIP(hipModuleGetFunction(&prog->f_cells, prog->module, "mk_cells"));
        MKH_HIP(hipModuleGetFunction(&prog->f_block, prog->module, "mk_block"));
        MKH_HIP(hipModuleGetFunction(&prog->f_grid, prog->module, "mk_grid"));
        MKH_HIP(hipModuleGetFunction(&prog->f_sheet, prog->module, "mk_sheet"));
        MKH_HIP(hipModuleGetFunction(&prog->f_chain, prog->module, "mk_chain"));
        MKH_HIP(hipModuleGetFunction(&prog->f_junction, prog->module,
                                     "mk_junction"));
        /* present only in kernels generated with gradients */
        if (hipModuleGetFunction(&prog->f_adj_step, prog->module,
                                 "mk_adj_step") != hipSuccess ||
            hipModuleGetFunction(&prog->f_adj_block, prog->module,
                                 "mk_adj_block") != hipSuccess ||
            hipModuleGetFunction(&prog->f_adj_probe, prog->module,
                                 "mk_adj_probe") != hipSuccess ||
            hipModuleGetFunction(&prog->f_adj_gather, prog->module,
                                 "mk_adj_gather") != hipSuccess) {
            prog->f_adj_step = prog->f_adj_block = prog->f_adj_probe =
                prog->f_adj_gather = nullptr;
            (void)hipGetLastError();
        }
        /* the persistent reverse sheet: stencil kernels whose tile fits */
        if (!prog->f_adj_step ||
            hipModuleGetFunction(&prog->f_adj_sheet, prog->module,
                                 "mk_adj_sheet") != hipSuccess) {
            prog->f_adj_sheet = nullptr;
            (void)hipGetLastError();
        }
        hipDeviceptr_t lp = nullptr;
        size_t lsz = 0;
        if (hipModuleGetGlobal(&lp, &lsz, prog->module, "mk_layout") ==
                hipSuccess && lsz == sizeof(prog->layout)) {
            MKH_HIP(hipMemcpyDtoH(prog->layout, lp, sizeof(prog->layout)));
            prog->have_layout = true;
        } else {
            (void)hipGetLastError();
        }
    }
    *program = prog.release();
    return 0;
}

int mkh_program_free(void* program)
{
    Program* p = (Program*)program;
    if (!p) return 0;
    if (p->module) (void)hipModuleUnload(p->module);
    delete p;
    return 0;
}

int mkh_sim_create(void* program, const mkh_sim_config* cfg, void** sim)
{
    if (!program || !cfg || !sim) return fail(MKH_E_ARG, "null argument");
    Program* p = (Program*)program;
    if (!p->module)
        return fail(MKH_E_STATE, "program was compiled without a device");
    if (cfg->nx < 1 || cfg->ny < 1 || cfg->n_state < 1)
        return fail(MKH_E_ARG, "invalid geometry");
    if (cfg->block_threads < 64 || cfg->block_threads > 1024 ||
        cfg->block_threads % 64)
        return fail(MKH_E_ARG, "block_threads must be a multiple of 64");
    if (cfg->precision != 32 && cfg->precision != 64)
        return fail(MKH_E_ARG, "precision must be 32 or 64");
    if (p->have_layout) {
        /* the configuration must describe the layout the kernel text was
         * generated for (a zero cells-per-thread turns that driver off) */
        const int* L = p->layout;
        const char* what = nullptr;
        int want = 0, got = 0;
        if (cfg->precision != L[0]) {
            what = "precision"; want = L[0]; got = cfg->precision;
        } else if (cfg->n_state != L[1]) {
            what = "n_state"; want = L[1]; got = cfg->n_state;
        } else if (cfg->n_field != L[2]) {
            what = "n_field"; want = L[2]; got = cfg->n_field;
        } else if (cfg->sheet_cpt > 0 && cfg->sheet_cpt != L[3]) {
            what = "sheet_cpt"; want = L[3]; got = cfg->sheet_cpt;
        } else if (cfg->chain_cpt > 0 && cfg->chain_cpt != L[4]) {
            what = "chain_cpt"; want = L[4]; got = cfg->chain_cpt;
        } else if (cfg->adj_cpt > 0 && cfg->adj_cpt != L[5]) {
            what = "adj_cpt"; want = L[5]; got = cfg->adj_cpt;
        }
        if (what)
            return fail(MKH_E_ARG, std::string(what) + " " +
                                       std::to_string(got) +
                                       " does not match the kernel's " +
                                       std::to_string(want));
    }
    SimBase* s;
    int r;
    if (cfg->precision == 64) {
        Sim<double>* d = new Sim<double>();
        r = d->create(p, cfg);
        s = d;
    } else {
        Sim<float>* d = new Sim<float>();
        r = d->create(p, cfg);
        s = d;
    }
    if (r) {
        delete s;
        return r;
    }
    *sim = s;
    return 0;
}

#define SIM ((SimBase*)sim)
#define CHECK_SIM() \
    if (!sim) return fail(MKH_E_ARG, "null simulation handle")

int mkh_sim_destroy(void* sim)
{
    if (sim) delete SIM;
    return 0;
}

int mkh_sim_set_state(void* sim, const double* s, long long n)
{
    CHECK_SIM();
    return SIM->set_state(s, n);
}

int mkh_sim_get_state(void* sim, double* s, long long n)
{
    CHECK_SIM();
    return SIM->get_state(s, n);
}

int mkh_sim_copy_state(void* dst, void* src)
{
    if (!dst || !src) return fail(MKH_E_ARG, "null simulation");
    return static_cast<SimBase*>(dst)->copy_state_from(
        static_cast<SimBase*>(src));
}

int mkh_ft_create(void* fiber, void* tissue, long long ctx, long long cty,
                  double gft, void** ft)
{
    if (!fiber || !tissue || !ft) return fail(MKH_E_ARG, "null argument");
    *ft = nullptr;
    SimBase* a = static_cast<SimBase*>(fiber);
    SimBase* b = static_cast<SimBase*>(tissue);
    if (a->cfg.precision != b->cfg.precision)
        return fail(MKH_E_ARG, "fiber and tissue differ in precision");
    int r;
    FtBase* h;
    if (a->cfg.precision == 64) {
        FtSim<double>* d = new FtSim<double>();
        r = d->create(static_cast<Sim<double>*>(a),
                      static_cast<Sim<double>*>(b), ctx, cty, gft);
        h = d;
    } else {
        FtSim<float>* d = new FtSim<float>();
        r = d->create(static_cast<Sim<float>*>(a),
                      static_cast<Sim<float>*>(b), ctx, cty, gft);
        h = d;
    }
    if (r) {
        delete h;
        return r;
    }
    *ft = h;
    return 0;
}

int mkh_ft_step(void* ft, double* t, int* status)
{
    if (!ft || !t || !status) return fail(MKH_E_ARG, "null argument");
    return static_cast<FtBase*>(ft)->step(t, status);
}

int mkh_ft_destroy(void* ft)
{
    delete static_cast<FtBase*>(ft);
    return 0;
}

int mkh_sim_set_fields(void* sim, const double* d, long long n)
{
    CHECK_SIM();
    return SIM->set_fields(d, n);
}

int mkh_sim_set_conductance(void* sim, double gx, double gy)
{
    CHECK_SIM();
    return SIM->set_conductance(gx, gy);
}

int mkh_sim_set_conductance_field(void* sim, const double* gx, long long ngx,
                                  const double* gy, long long ngy)
{
    CHECK_SIM();
    return SIM->set_conductance_field(gx, ngx, gy, ngy);
}

int mkh_sim_set_connections(void* sim, const long long* i,
                            const long long* j, const double* g, long long n)
{
    CHECK_SIM();
    return SIM->set_connections(i, j, g, n);
}

int mkh_sim_set_domains(void* sim, const long long* offsets, long long nd)
{
    CHECK_SIM();
    return SIM->set_domains(offsets, nd);
}

int mkh_sim_set_pacing_all(void* sim)
{
    CHECK_SIM();
    return SIM->set_pacing(0, nullptr, 0);
}

int mkh_sim_set_pacing_none(void* sim)
{
    CHECK_SIM();
    return SIM->set_pacing(3, nullptr, 0);
}

int mkh_sim_set_pacing_rect(void* sim, long long nx, long long ny,
                            long long x, long long y)
{
    CHECK_SIM();
    long long r[4] = {nx, ny, x, y};
    return SIM->set_pacing(1, r, 4);
}

int mkh_sim_set_pacing_list(void* sim, const long long* cells, long long n)
{
    CHECK_SIM();
    return SIM->set_pacing(2, cells, n);
}

int mkh_sim_set_logging(void* sim, const int* colmap, long long nq, int nc,
                        int nan_check)
{
    CHECK_SIM();
    return SIM->set_logging(colmap, nq, nc, nan_check);
}

int mkh_sim_set_nan_cell(void* sim, long long cell)
{
    CHECK_SIM();
    return SIM->set_nan_cell(cell);
}

int mkh_sim_init(void* sim, double tmin, double tmax, double dt,
                 const double* events, long long n_events,
                 double log_interval, long long* n_log_rows)
{
    CHECK_SIM();
    return SIM->init(tmin, tmax, dt, events, n_events, log_interval,
                     n_log_rows);
}

int mkh_sim_set_log(void* sim, void* values, double* times, double* paces)
{
    CHECK_SIM();
    return SIM->set_log(values, times, paces);
}

int mkh_sim_step(void* sim, double* time, int* status)
{
    CHECK_SIM();
    return SIM->step(time, status);
}

int mkh_sim_clean(void* sim)
{
    CHECK_SIM();
    return SIM->clean();
}

int mkh_sim_set_option(void* sim, const char* key, long long value)
{
    CHECK_SIM();
    if (!key) return fail(MKH_E_ARG, "null option name");
    return SIM->set_option(key, value);
}

int mkh_sim_get_timing(void* sim, double* kernel_ms, long long* launches,
                       long long* cell_steps)
{
    CHECK_SIM();
    return SIM->timing(kernel_ms, launches, cell_steps);
}

int mkh_sim_reset_timing(void* sim)
{
    CHECK_SIM();
    return SIM->reset_timing();
}

int mkh_adj_run(void* sim, double tmin, double tmax, double dt,
                const double* events, long long n_events,
                const long long* probes, long long n_probes,
                const double* target, long long n_target_rows,
                int checkpoint_every, double* loss, double* grad_fields,
                double* grad_g)
{
    CHECK_SIM();
    return SIM->adj_run(tmin, tmax, dt, events, n_events, probes, n_probes,
                        target, n_target_rows, checkpoint_every, loss,
                        grad_fields, grad_g, nullptr);
}

int mkh_host_alloc(long long bytes, void** ptr)
{
    if (!ptr || bytes < 0) return fail(MKH_E_ARG, "bad argument");
    *ptr = nullptr;
    if (bytes == 0) return 0;
    MKH_HIP(hipHostMalloc(ptr, (size_t)bytes, hipHostMallocDefault));
    return 0;
}

int mkh_host_free(void* ptr)
{
    if (ptr) MKH_HIP(hipHostFree(ptr));
    return 0;
}

int mkh_adj_timing(void* sim, double* ms3)
{
    CHECK_SIM();
    if (!ms3) return fail(MKH_E_ARG, "null argument");
    return SIM->adj_timing(ms3);
}

int mkh_adj_reverse(void* sim, int* kind, long long* launches)
{
    CHECK_SIM();
    if (!kind || !launches) return fail(MKH_E_ARG, "null argument");
    return SIM->adj_reverse(kind, launches);
}

int mkh_adj_set_cell_order(void* sim, const long long* order,
                           long long n_cells)
{
    CHECK_SIM();
    if (n_cells < 0) return fail(MKH_E_ARG, "bad argument");
    return SIM->set_cell_order(order, n_cells);
}

int mkh_adj_trace(void* sim, double tmin, double tmax, double dt,
                  const double* events, long long n_events,
                  const long long* probes, long long n_probes, double* trace,
                  long long n_rows)
{
    if (!sim || !trace) return fail(MKH_E_ARG, "null argument");
    return SIM->adj_run(tmin, tmax, dt, events, n_events, probes, n_probes,
                        nullptr, n_rows, 1, nullptr, nullptr, nullptr, trace);
}

int mkh_adj_trace_static(void* sim, double tmin, double tmax, double dt,
                         const double* events, long long n_events,
                         const long long* probes, long long n_probes,
                         double* trace, long long n_rows)
{
    if (!sim || !trace) return fail(MKH_E_ARG, "null argument");
    return SIM->trace_static(tmin, tmax, dt, events, n_events, probes,
                             n_probes, trace, n_rows);
}

int mkh_sim_driver(void* sim, int* driver)
{
    CHECK_SIM();
    if (driver) *driver = SIM->get_driver();
    return 0;
}

int mkh_sim_log_rows(void* sim, long long* rows)
{
    CHECK_SIM();
    if (rows) *rows = SIM->log_rows();
    return 0;
}

int mkh_pacing_levels(const double* events, long long n_events, double t0,
                      const double* times, long long n_times, double* levels,
                      double* next_times)
{
    MkESys es;
    int f = es.init(t0, events, n_events);
    if (f) return pacing_error(f);
    for (long long k = 0; k < n_times; ++k) {
        f = es.advance(times[k]);
        if (f) return pacing_error(f);
        if (levels) levels[k] = es.level();
        if (next_times) next_times[k] = es.next_time();
    }
    return 0;
}

int mkh_plan(double tmin, double tmax, double dt, const double* events,
             long long n_events, double log_interval, double* t,
             double* step, double* pace, int* log, long long max_steps,
             long long* n_steps)
{
    Planner p;
    int f = p.init(tmin, tmax, dt, events, n_events, log_interval);
    if (f) return pacing_error(f);
    long long k = 0;
    StepRec r;
    while (!p.done && k < max_steps) {
        f = p.next(r);
        if (f) return pacing_error(f);
        if (t) t[k] = r.t;
        if (step) step[k] = r.dt;
        if (pace) pace[k] = r.pace;
        if (log) log[k] = r.log ? 1 : 0;
        ++k;
    }
    if (n_steps) *n_steps = k;
    return 0;
}

}  // extern "C"
