/*
 * mk_drivers.hip -- execution drivers for the generated cell step.
 *
 * This text is appended to a generated kernel source after the per-model
 * function
 *     template <bool LOG> __device__ void mk_cell(Real* y, Real t, Real dt,
 *         Real pace, Real idiff, const Real* fv, Real* q);
 * and the macros MK_NS (states), MK_NF (fields), MK_NQ (logged per-cell
 * quantities), MK_IVM (membrane potential index), MK_COUPLING (0 none,
 * 1 stencil, 2 conductance field, 3 graph), MK_BT (threads per workgroup)
 * and MK_BCELLS (LDS capacity of the block driver, in cells).
 *
 * One time step of one cell is (openclsim.c:1066-1096, openclsim.cl:298-367):
 *   idiff = coupling current from the membrane potentials at time t,
 *   pace  = the protocol level if the cell is paced,
 *   y     = FE / RL update of the state from t to t + dt,
 * with the state at t, the intermediaries and idiff written to the log at
 * log steps. The reference runs this as 2-3 kernels per step with the state
 * in AoS layout; here the coupling and the update are fused in one pass over
 * SoA state, and five drivers cover the geometries:
 *
 *   mk_cells  uncoupled cells: one thread per cell keeps its state in
 *             registers across every step of the launch (no synchronisation);
 *   mk_block  one workgroup per independent domain (a cable, a small sheet,
 *             a connected component of a graph): V is exchanged through LDS
 *             with one barrier per step, all steps of the launch;
 *   mk_grid   any size: one fused step per launch, V of neighbours read from
 *             the previous state buffer, new state written to the other
 *             buffer (ping-pong);
 *   mk_sheet  large homogeneous 2D sheets: a persistent cooperative launch,
 *             one 64 x 16 tile per workgroup with its state in registers
 *             for every step of the launch; tile edges are exchanged
 *             through L2 as step-tagged words;
 *   mk_chain  graphs of disjoint chains (cables): the same persistent
 *             scheme over the cells laid end to end, a cable spread over
 *             as many tiles as its length needs.
 *
 * Halting: at a log step, the thread of the caller's cell 0 (nan_cell in
 * the device's cell order) records the log row in halt[0] if state 0 is NaN
 * (openclsim.c:1087-1089). The host ends launches
 * at log steps, and a launch returns at once if halt[0] holds a row other
 * than its own, so steps after the halting step are never taken.
 */
typedef MkArgsT<Real> MkArgs;
typedef MkStepT<Real> MkStep;

/* The layout this text was generated for: precision (bits), states,
 * fields, and the cells per thread of the sheet, chain and reverse-sheet
 * tiles. The host library reads it when it loads the module and rejects a
 * simulation configured for another layout. */
extern "C" __device__ const int mk_layout[6] = {
    (int)(sizeof(Real) * 8), MK_NS, MK_NF, MK_SCPT, MK_CCPT, MK_ACPT};

#define MK_NFA (MK_NF > 0 ? MK_NF : 1)
#define MK_NQA (MK_NQ > 0 ? MK_NQ : 1)

/* The helpers below take cell indices of type I: long long in the forward
 * drivers, int in the reverse kernels (at most 2^31 - 1 cells), where the
 * 64-bit division and address arithmetic would be a fifth of the work. */
template <typename I>
__device__ __forceinline__ bool mk_is_paced(const MkArgs& a, I c, I ix, I iy)
{
    if (a.pace_mode == 0) return true;
    if (a.pace_mode == 1)
        return ix >= (I)a.px0 && ix < (I)a.px1 && iy >= (I)a.py0
               && iy < (I)a.py1;
    if (a.pace_mode == 2) return a.pace_mask[c] != 0;
    return false;
}

/* Homogeneous zero-flux stencil (diff_step, openclsim.cl:384-435); V is
 * indexed by cell id, same operation order as the reference. */
template <typename I>
__device__ __forceinline__ Real mk_stencil(
    const Real* __restrict__ V, I c, I ix, I iy, I nx, I ny, Real gx, Real gy)
{
    const Real v = V[c];
    Real i;
    if (nx > 1) {
        if (ix == 0) i = gx * (v - V[c + 1]);
        else if (ix == nx - 1) i = gx * (v - V[c - 1]);
        else i = gx * (((Real)2 * v - V[c - 1]) - V[c + 1]);
    } else {
        i = (Real)0;
    }
    if (ny > 1) {
        if (iy == 0) i += gy * (v - V[c + nx]);
        else if (iy == ny - 1) i += gy * (v - V[c - nx]);
        else i += gy * (((Real)2 * v - V[c - nx]) - V[c + nx]);
    }
    return i;
}

/* Conductance-field stencil (diff_hetero, openclsim.cl:452-487). */
template <typename I>
__device__ __forceinline__ Real mk_hetero(
    const Real* __restrict__ V, const Real* __restrict__ gxf,
    const Real* __restrict__ gyf, I c, I ix, I iy, I nx, I ny)
{
    const Real v = V[c];
    Real i = (Real)0;
    if (nx > 1) {
        if (ix > 0) i += gxf[c - iy - 1] * (v - V[c - 1]);
        if (ix < nx - 1) i += gxf[c - iy] * (v - V[c + 1]);
    }
    if (ny > 1) {
        if (iy > 0) i += gyf[c - nx] * (v - V[c - nx]);
        if (iy < ny - 1) i += gyf[c] * (v - V[c + nx]);
    }
    return i;
}

/* Graph coupling as a gather over the cell's connections, in the order of
 * the connection list: the same additions, in the same order, as the
 * reference's edge scatter (diff_arb_step, openclsim.cl:537-556) when its
 * atomics happen to run in list order. V is indexed by `base + j`. With
 * copies > 1 (a graph of identical copies held interleaved, MkArgs.gcopies)
 * the rows are those of one copy: cell `self` reads row self / copies, and
 * a neighbour position q is cell q * copies + self % copies. */
template <typename I>
__device__ __forceinline__ Real mk_graph(
    const Real* __restrict__ V, const int* __restrict__ ptr,
    const int* __restrict__ idx, const Real* __restrict__ g, I row, I self,
    int copies = 1)
{
    const Real v = V[self];
    Real i = (Real)0;
    const bool cp = copies > 1;
    I lane = 0;
    if (cp) {
        row = self / copies;
        lane = self - row * copies;
    }
    int e = ptr[row];
    const int e1 = ptr[row + 1];
    /* four connections per round: their index, conductance and neighbour
     * loads are in flight together (one dependent load chain per round
     * instead of per connection); the last round is masked; the additions
     * keep the list order */
    for (; e < e1; e += 4) {
        int j[4];
        Real gg[4], vj[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) j[k] = e + k < e1 ? idx[e + k] : 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (e + k < e1) {
                const int q = j[k] >= 0 ? j[k] : ~j[k];
                gg[k] = g[e + k];
                vj[k] = V[cp ? (I)q * copies + lane : (I)q];
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (e + k < e1) {
                if (j[k] >= 0) i = i + gg[k] * (v - vj[k]);
                else i = i - gg[k] * (vj[k] - v);
            }
        }
    }
    return i;
}

__device__ __forceinline__ void mk_write_log(
    const MkArgs& a, long long c, int row, const Real* q)
{
    for (int k = 0; k < MK_NQ; ++k) {
        const int col = a.colmap[(long long)k * a.npad + c];
        if (col >= 0) a.logbuf[(long long)row * a.ncols + col] = q[k];
    }
}

/* The logged quantities set by part P of the split cell step (mk_cell_a:
 * 0, mk_cell_b: 1). */
template <int P>
__device__ __forceinline__ void mk_write_log_part(
    const MkArgs& a, long long c, int row, const Real* q)
{
#pragma unroll
    for (int k = 0; k < MK_NQ; ++k) {
        if (mk_qpart[k] != P) continue;
        const int col = a.colmap[(long long)k * a.npad + c];
        if (col >= 0) a.logbuf[(long long)row * a.ncols + col] = q[k];
    }
}

/* Adjoint support: the state after local step i of a launch goes to
 * traj + i * n_state * npad (SoA), the trajectory the reverse pass reads. */
__device__ __forceinline__ void mk_store_traj(const MkArgs& a, Real* traj,
                                              int i, long long c,
                                              const Real* y)
{
    Real* __restrict__ tr = traj + (long long)i * MK_NS * a.npad + c;
#pragma unroll
    for (int st = 0; st < MK_NS; ++st) tr[(long long)st * a.npad] = y[st];
}

__device__ __forceinline__ bool mk_halted(const MkArgs& a, int own_row)
{
    const int h = __atomic_load_n(a.halt, __ATOMIC_RELAXED);
    return h >= 0 && h != own_row;
}

/* ------------------------------------------------------------------------
 * Uncoupled cells: one thread per cell, state in registers for all steps.
 * ---------------------------------------------------------------------- */
extern "C" __global__ void __launch_bounds__(MK_BT)
mk_cells(const MkArgs a, const int s0, const int s1, Real* const traj)
{
    if (mk_halted(a, a.steps[s1 - 1].row)) return;
    const long long c = (long long)blockIdx.x * MK_BT + threadIdx.x;
    if (c >= a.n) return;
    Real y[MK_NS];
    Real fv[MK_NFA];
    Real q[MK_NQA];
#pragma unroll
    for (int i = 0; i < MK_NS; ++i) y[i] = a.state_a[(long long)i * a.npad + c];
#pragma unroll
    for (int k = 0; k < MK_NF; ++k) fv[k] = a.fields[(long long)k * a.npad + c];
    const long long ix = c % a.nx, iy = c / a.nx;
    const bool paced = mk_is_paced(a, c, ix, iy);
    for (int s = s0; s < s1; ++s) {
        const MkStep p = a.steps[s];
        const Real pace = paced ? p.pace : (Real)0;
        if (p.row >= 0 && a.nan_check && c == a.nan_cell && y[0] != y[0])
            __atomic_store_n(a.halt, p.row, __ATOMIC_RELAXED);
        /* One instance of the cell code: logged values are register
         * copies, written out only at log steps */
        mk_cell<true>(y, p.t, p.dt, pace, (Real)0, fv, q);
        if (p.row >= 0) mk_write_log(a, c, p.row, q);
        if (traj) mk_store_traj(a, traj, s - s0, c, y);
    }
#pragma unroll
    for (int i = 0; i < MK_NS; ++i) a.state_a[(long long)i * a.npad + c] = y[i];
}

/* ------------------------------------------------------------------------
 * One workgroup per domain; V exchanged through LDS, one barrier per step.
 * ---------------------------------------------------------------------- */
extern "C" __global__ void __launch_bounds__(MK_BT)
mk_block(const MkArgs a, const int s0, const int s1, Real* const traj)
{
    __shared__ Real vsh[2][MK_BCELLS];
    const bool stop = mk_halted(a, a.steps[s1 - 1].row);
    __syncthreads();
    if (stop) return;
    const long long off = a.dom_off[blockIdx.x];
    const int len = (int)(a.dom_off[blockIdx.x + 1] - off);
    /* A single stencil domain is the whole (nx, ny) grid; graph domains are
     * addressed through dom_cells and a local CSR. */
    const long long nx = a.nx, ny = a.ny;
    for (int li = threadIdx.x; li < len; li += MK_BT) {
        const long long c = a.dom_cells ? a.dom_cells[off + li] : off + li;
        vsh[0][li] = a.state_a[(long long)MK_IVM * a.npad + c];
    }
    __syncthreads();
    int cur = 0;
    Real y[MK_NS];
    Real fv[MK_NFA];
    Real q[MK_NQA];
    for (int s = s0; s < s1; ++s) {
        const MkStep p = a.steps[s];
        for (int li = threadIdx.x; li < len; li += MK_BT) {
            const long long c = a.dom_cells ? a.dom_cells[off + li] : off + li;
#pragma unroll
            for (int i = 0; i < MK_NS; ++i)
                y[i] = a.state_a[(long long)i * a.npad + c];
#pragma unroll
            for (int k = 0; k < MK_NF; ++k)
                fv[k] = a.fields[(long long)k * a.npad + c];
            const long long ix = c % nx, iy = c / nx;
#if MK_COUPLING == 1
            const Real idiff = mk_stencil<long long>(vsh[cur], li, ix, iy, nx,
                                                     ny, a.gx, a.gy);
#elif MK_COUPLING == 2
            const Real idiff = mk_hetero<long long>(vsh[cur], a.gx_field,
                                                    a.gy_field, li, ix, iy,
                                                    nx, ny);
#elif MK_COUPLING == 3
            const Real idiff = mk_graph<long long>(vsh[cur], a.csr_ptr,
                                                   a.csr_idx, a.csr_g,
                                                   off + li, li);
#else
            const Real idiff = (Real)0;
#endif
            const Real pace = mk_is_paced(a, c, ix, iy) ? p.pace : (Real)0;
            if (p.row >= 0 && a.nan_check && c == a.nan_cell && y[0] != y[0])
                __atomic_store_n(a.halt, p.row, __ATOMIC_RELAXED);
            /* One instance of the cell code: logged values are register
             * copies, written out only at log steps */
            mk_cell<true>(y, p.t, p.dt, pace, idiff, fv, q);
            if (p.row >= 0) mk_write_log(a, c, p.row, q);
#pragma unroll
            for (int i = 0; i < MK_NS; ++i)
                a.state_a[(long long)i * a.npad + c] = y[i];
            if (traj) mk_store_traj(a, traj, s - s0, c, y);
            vsh[cur ^ 1][li] = y[MK_IVM];
        }
        cur ^= 1;
        __syncthreads();
    }
}

/* ------------------------------------------------------------------------
 * Whole grid, one fused step per launch, ping-pong state buffers.
 * 2D grids are tiled 64 x (MK_BT / 64); workgroups are renumbered so that
 * each of the 8 XCDs (which receive workgroups round-robin) gets one
 * contiguous band of tiles, keeping the neighbour rows in its own L2.
 * ---------------------------------------------------------------------- */
__device__ __forceinline__ long long mk_xcd_block(long long b, long long nb)
{
    const long long q = nb / 8, r = nb % 8;
    const long long x = b % 8, k = b / 8;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

#ifndef MK_GBLOCKS
#define MK_GBLOCKS 0
#endif
#if MK_GBLOCKS > 0
#define MK_GRID_BOUNDS __launch_bounds__(MK_BT, MK_GBLOCKS)
#else
#define MK_GRID_BOUNDS __launch_bounds__(MK_BT)
#endif
extern "C" __global__ void MK_GRID_BOUNDS
mk_grid(const MkArgs a, const Real* __restrict__ src, Real* __restrict__ dst,
        const int s)
{
    const MkStep p = a.steps[s];
    if (mk_halted(a, p.row)) return;
    const long long b = mk_xcd_block(blockIdx.x, gridDim.x);
    long long c, ix, iy;
#if MK_COUPLING == 1 || MK_COUPLING == 2
    if (a.ny > 1) {
        const long long tiles_x = (a.nx + 63) / 64;
        ix = (b % tiles_x) * 64 + (threadIdx.x & 63);
        iy = (b / tiles_x) * (MK_BT / 64) + (threadIdx.x >> 6);
        if (ix >= a.nx || iy >= a.ny) return;
        c = ix + iy * a.nx;
    } else {
        c = b * MK_BT + threadIdx.x;
        if (c >= a.n) return;
        ix = c;
        iy = 0;
    }
#else
    c = b * MK_BT + threadIdx.x;
    if (c >= a.n) return;
    ix = c % a.nx;
    iy = c / a.nx;
#endif
    Real y[MK_NS];
    Real fv[MK_NFA];
    Real q[MK_NQA];
#pragma unroll
    for (int i = 0; i < MK_NS; ++i) y[i] = src[(long long)i * a.npad + c];
#pragma unroll
    for (int k = 0; k < MK_NF; ++k) fv[k] = a.fields[(long long)k * a.npad + c];
    const Real* __restrict__ V = src + (long long)MK_IVM * a.npad;
#if MK_COUPLING == 1
    Real idiff = mk_stencil(V, c, ix, iy, a.nx, a.ny, a.gx, a.gy);
#elif MK_COUPLING == 2
    Real idiff = mk_hetero(V, a.gx_field, a.gy_field, c, ix, iy,
                                 a.nx, a.ny);
#elif MK_COUPLING == 3
    Real idiff = mk_graph<long long>(V, a.csr_ptr, a.csr_idx, a.csr_g,
                                           c, c, a.gcopies);
#else
    Real idiff = (Real)0;
#endif
    /* junction current of the fiber-tissue driver, added after the stencil
     * as the reference adds it (diff_step_fiber_tissue, openclsim.cl:601-
     * 628); zero elsewhere, where adding nothing keeps idiff bit for bit */
    if (a.idiff_extra) {
        const Real e = a.idiff_extra[c];
        if (e != (Real)0) idiff += e;
    }
    const Real pace = mk_is_paced(a, c, ix, iy) ? p.pace : (Real)0;
    if (p.row >= 0 && a.nan_check && c == a.nan_cell && y[0] != y[0])
        __atomic_store_n(a.halt, p.row, __ATOMIC_RELAXED);
    /* One instance of the cell code: logged values are register
     * copies, written out only at log steps */
    mk_cell<true>(y, p.t, p.dt, pace, idiff, fv, q);
    if (p.row >= 0) mk_write_log(a, c, p.row, q);
#pragma unroll
    for (int i = 0; i < MK_NS; ++i) dst[(long long)i * a.npad + c] = y[i];
}

/* ------------------------------------------------------------------------
 * Persistent 2D sheet: register-resident state, LDS halos, neighbour sync.
 *
 * Each workgroup (256 threads = 4 waves) owns a tile of 64 x MK_STH cells
 * (MK_STH = 4 * MK_SCPT) for every step of the launch; thread (lx, w) keeps
 * the full state of cells (x0 + lx, y0 + w * MK_SCPT + k), k < MK_SCPT, in
 * registers. The tile's membrane potentials, plus a one-cell halo ring, live
 * in LDS (double-buffered by step). After each step a tile writes its four
 * edge rows/columns of V to global memory as step-tagged words; in the next
 * step each thread of its halo ring waits for its neighbour value of that
 * step and copies it into LDS (one barrier per step). No state goes to HBM
 * inside a launch: per cell-step the only global traffic is the edge
 * exchange.
 *
 * A step runs the generated cell step in two parts (mk_cell_a, mk_cell_b):
 * part A of all the thread's cells first, everything that does not depend
 * on the diffusion current (for LR1991 all but the update of V), with what
 * part B reads kept in LDS; then the halo wait and the barrier; then each
 * cell's diffusion current and part B. The neighbours' edge words travel
 * while part A computes, so the wait rarely finds them missing. The
 * operations, and so the results, are those of mk_cell.
 *
 * The launch is cooperative, so all tiles are resident at once. Every wait
 * is bounded (spin_limit polls); on a timeout the tile records an error and
 * returns, its neighbours then time out too, and every wave exits.
 * ---------------------------------------------------------------------- */
#ifndef MK_SCPT
#define MK_SCPT 4
#endif
#ifndef MK_SWAVES
#define MK_SWAVES 4     /* waves per SIMD the register budget must allow */
#endif
#define MK_STH (4 * MK_SCPT)
typedef MkSheetT<Real> MkSheet;

/* Exchange between tiles. A value crosses as one 64-bit word per 32 bits
 * (one in single precision, two in double), each word holding the step the
 * value belongs to (high half) and 32 bits of the value (low half), written
 * and read with agent-scope atomic accesses of the whole word. A reader
 * polls the word until its step is the one it needs: an atomic access of
 * one location sees one whole write of it, so no ordering between
 * different locations (a separate flag, a fence) is needed, and a writer
 * never waits for its stores. Atomics at agent scope are coherent across
 * the XCDs' separate L2 caches. Each value has two slots, by step parity:
 * a writer reuses a slot two steps later, which needs its neighbour's value
 * of the step in between, which the neighbour writes only after it has
 * read the slot. */
typedef unsigned long long MkXWord;
#define MK_XW ((int)(sizeof(Real) / 4))

__device__ __forceinline__ void mk_tput(MkXWord* p, Real v, int step)
{
    const MkXWord t = (MkXWord)(unsigned)step << 32;
    if constexpr (sizeof(Real) == 4) {
        __hip_atomic_store(p, t | (MkXWord)__float_as_uint((float)v),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        const MkXWord b = (MkXWord)__double_as_longlong((double)v);
        __hip_atomic_store(p, t | (b >> 32), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(p + 1, t | (b & 0xffffffffull), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
}

/* Polls the value at p until it belongs to `step` (or a later one);
 * returns false after spin_limit polls. */
__device__ __forceinline__ bool mk_tget(const MkXWord* p, int step,
                                        int spin_limit, Real& v)
{
    MkXWord a = 0, b = 0;
    for (int spins = 0;; ++spins) {
        a = __hip_atomic_load(const_cast<MkXWord*>(p), __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
        bool ready = (int)(a >> 32) >= step;
        if constexpr (sizeof(Real) == 8) {
            b = __hip_atomic_load(const_cast<MkXWord*>(p + 1),
                                  __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ready = ready && (int)(b >> 32) >= step;
        }
        if (ready) break;
        if (spins >= spin_limit) return false;
        __builtin_amdgcn_s_sleep(1);
    }
    if constexpr (sizeof(Real) == 4) {
        v = (Real)__uint_as_float((unsigned)a);
    } else {
        v = (Real)__longlong_as_double(
            (long long)(((a & 0xffffffffull) << 32) | (b & 0xffffffffull)));
    }
    return true;
}

/* Value k of edge e (0 bottom row, 1 top row, 2 left column, 3 right
 * column) of a tile, slot of step parity par. */
__device__ __forceinline__ MkXWord* mk_edge(const MkSheet& sh, int par,
                                            int tile, int e, int k)
{
    return sh.edges + (((long long)(par * sh.tiles_x * sh.tiles_y + tile) * 4
                        + e) * 64 + k) * MK_XW;
}

extern "C" __global__ void __launch_bounds__(256, MK_SWAVES)
mk_sheet(const MkArgs a, const MkSheet sh, const int s0, const int s1)
{
    __shared__ Real vt[2][MK_STH + 2][66];
    /* what part B of each cell's step reads of part A's values: in LDS
     * rather than registers, where it would add to the live state of all
     * the thread's cells across the halo wait */
    __shared__ Real ca[MK_NCARRY][MK_SCPT][256];
    __shared__ int wg_flag;
    if (threadIdx.x == 0)
        wg_flag = mk_halted(a, a.steps[s1 - 1].row) ? 1 : 0;
    __syncthreads();
    if (wg_flag) return;

    const int tile = blockIdx.x;
    const int tx = tile % sh.tiles_x, ty = tile / sh.tiles_x;
    const int lx = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long gx0 = (long long)tx * 64, gy0 = (long long)ty * MK_STH;
    const long long ix = gx0 + lx;
    const long long nx = a.nx, ny = a.ny;

    Real y[MK_SCPT][MK_NS];
    Real fv[MK_SCPT][MK_NFA];
    Real q[MK_NQA];
    bool act[MK_SCPT];
    bool paced[MK_SCPT];
    /* cell indices in 32 bits (a co-resident sheet has far fewer than 2^31
     * cells) and whether the NaN test reads the cell */
    int cid[MK_SCPT];
    bool nanc[MK_SCPT];
#pragma unroll
    for (int k = 0; k < MK_SCPT; ++k) {
        const int r = w * MK_SCPT + k;
        const long long iy = gy0 + r;
        act[k] = ix < nx && iy < ny;
        const long long c = act[k] ? ix + iy * nx : 0;
        cid[k] = (int)c;
        nanc[k] = act[k] && a.nan_check && c == a.nan_cell;
        paced[k] = act[k] && mk_is_paced(a, c, ix, iy);
#pragma unroll
        for (int i = 0; i < MK_NS; ++i)
            y[k][i] = act[k] ? a.state_a[(long long)i * a.npad + c] : (Real)0;
#pragma unroll
        for (int f = 0; f < MK_NF; ++f)
            fv[k][f] = act[k] ? a.fields[(long long)f * a.npad + c] : (Real)0;
        vt[0][r + 1][lx + 1] = y[k][MK_IVM];
    }
    /* Halo ring at the start of the launch, from the global state */
    {
        const Real* V = a.state_a + (long long)MK_IVM * a.npad;
        const int t = threadIdx.x;
        long long hx = -1, hy = -1;
        int lr = 0, lc = 0;
        if (t < 66) {
            hx = gx0 - 1 + t; hy = gy0 - 1; lr = 0; lc = t;
        } else if (t < 132) {
            hx = gx0 - 1 + (t - 66); hy = gy0 + MK_STH; lr = MK_STH + 1;
            lc = t - 66;
        } else if (t < 132 + MK_STH) {
            hx = gx0 - 1; hy = gy0 + (t - 132); lr = t - 132 + 1; lc = 0;
        } else if (t < 132 + 2 * MK_STH) {
            hx = gx0 + 64; hy = gy0 + (t - 132 - MK_STH);
            lr = t - 132 - MK_STH + 1; lc = 65;
        }
        if (hx >= 0 && hy >= 0 && hx < nx && hy < ny)
            vt[0][lr][lc] = V[hx + hy * nx];
    }
    __syncthreads();

    const int ntiles = sh.tiles_x * sh.tiles_y;
    const int nb_tile[4] = {ty > 0 ? tile - sh.tiles_x : -1,
                            ty + 1 < sh.tiles_y ? tile + sh.tiles_x : -1,
                            tx > 0 ? tile - 1 : -1,
                            tx + 1 < sh.tiles_x ? tile + 1 : -1};
    (void)ntiles;
    int cur = 0;
    for (int s = s0; s < s1; ++s) {
        const int i = s - s0;
        const MkStep p = a.steps[s];
        const bool last = s + 1 == s1;
        const int par = (sh.g0 + i) & 1;
        /* Part A of every cell's step: all that does not depend on the
         * diffusion current, so the neighbours' edges of the previous step
         * are on their way meanwhile */
#pragma unroll
        for (int k = 0; k < MK_SCPT; ++k) {
            if (act[k]) {
                if (p.row >= 0 && nanc[k] && y[k][0] != y[k][0])
                    __atomic_store_n(a.halt, p.row, __ATOMIC_RELAXED);
                const Real pace = paced[k] ? p.pace : (Real)0;
                Real cy[MK_NCARRY];
                /* logged values are register copies, written out only at
                 * log steps */
                mk_cell_a<true>(y[k], p.t, p.dt, pace, fv[k], q, cy);
#pragma unroll
                for (int j = 0; j < MK_NCARRY; ++j)
                    ca[j][k][threadIdx.x] = cy[j];
                if (p.row >= 0) mk_write_log_part<0>(a, cid[k], p.row, q);
            }
            /* one cell after the other: interleaving the cells' part A
             * buys no latency the other waves do not hide, and multiplies
             * the live registers */
            __builtin_amdgcn_sched_barrier(0);
        }
        if (i > 0) {
            /* The neighbours' edges of the previous step (south: its top
             * row, north: its bottom row, west: its right column, east: its
             * left column) into the halo ring, each halo thread waiting for
             * its own value */
            const int step = sh.g0 + i;
            const int hpar = (step - 1) & 1;
            const int t = threadIdx.x;
            int nbt = -1, e = 0, k = 0, lr = 0, lc = 0;
            if (t < 64) {
                nbt = nb_tile[0]; e = 1; k = t; lr = 0; lc = t + 1;
            } else if (t < 128) {
                nbt = nb_tile[1]; e = 0; k = t - 64; lr = MK_STH + 1;
                lc = t - 63;
            } else if (t < 128 + MK_STH) {
                nbt = nb_tile[2]; e = 3; k = t - 128; lr = t - 127; lc = 0;
            } else if (t < 128 + 2 * MK_STH) {
                nbt = nb_tile[3]; e = 2; k = t - 128 - MK_STH;
                lr = t - 127 - MK_STH; lc = 65;
            }
            /* only cells of the sheet are written: the columns of a partial
             * last tile column, the rows of a partial last tile row, are
             * not */
            if (nbt >= 0 && (e < 2 ? gx0 + k < nx : gy0 + k < ny)) {
                Real v;
                if (mk_tget(mk_edge(sh, hpar, nbt, e, k), step,
                            sh.spin_limit, v)) {
                    vt[cur][lr][lc] = v;
                } else {
                    __hip_atomic_store(sh.err, 1, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    wg_flag = 1;
                }
            }
            __syncthreads();
            if (wg_flag) return;
        }
        /* Part B: the diffusion current from the potentials at t, then the
         * rest of the step */
#pragma unroll
        for (int k = 0; k < MK_SCPT; ++k) {
            const int r = w * MK_SCPT + k;
            if (act[k]) {
                const long long iy = gy0 + r;
                const Real v = vt[cur][r + 1][lx + 1];
                Real id;
                if (nx > 1) {
                    if (ix == 0) id = a.gx * (v - vt[cur][r + 1][lx + 2]);
                    else if (ix == nx - 1) id = a.gx * (v - vt[cur][r + 1][lx]);
                    else id = a.gx * (((Real)2 * v - vt[cur][r + 1][lx])
                                      - vt[cur][r + 1][lx + 2]);
                } else {
                    id = (Real)0;
                }
                if (ny > 1) {
                    if (iy == 0) id += a.gy * (v - vt[cur][r + 2][lx + 1]);
                    else if (iy == ny - 1) id += a.gy * (v - vt[cur][r][lx + 1]);
                    else id += a.gy * (((Real)2 * v - vt[cur][r][lx + 1])
                                       - vt[cur][r + 2][lx + 1]);
                }
                const Real pace = paced[k] ? p.pace : (Real)0;
                const long long c = cid[k];
                Real cy[MK_NCARRY];
#pragma unroll
                for (int j = 0; j < MK_NCARRY; ++j)
                    cy[j] = ca[j][k][threadIdx.x];
                mk_cell_b<true>(y[k], p.t, p.dt, pace, id, fv[k], q, cy);
                if (p.row >= 0) mk_write_log_part<1>(a, c, p.row, q);
                if (sh.traj) mk_store_traj(a, sh.traj, i, c, y[k]);
                vt[cur ^ 1][r + 1][lx + 1] = y[k][MK_IVM];
            }
        }
        /* The tile's edges of this step for its neighbours: the bottom row
         * (wave 0's first cells), the top row (wave 3's last), and the left
         * and right columns (lanes 0 and 63 of every wave, their cells one
         * after the other) */
        if (!last) {
            const int step = sh.g0 + i + 1;
            if (w == 0 && act[0])
                mk_tput(mk_edge(sh, par, tile, 0, lx), y[0][MK_IVM], step);
            if (w == 3 && act[MK_SCPT - 1])
                mk_tput(mk_edge(sh, par, tile, 1, lx),
                        y[MK_SCPT - 1][MK_IVM], step);
            if (lx == 0 || lx == 63) {
                MkXWord* ep = mk_edge(sh, par, tile, lx == 0 ? 2 : 3,
                                      w * MK_SCPT);
#pragma unroll
                for (int k = 0; k < MK_SCPT; ++k)
                    if (act[k]) mk_tput(ep + k * MK_XW, y[k][MK_IVM], step);
            }
        }
        /* The next step's barrier (after its halo) orders this step's LDS
         * writes before its reads */
        cur ^= 1;
    }
    /* The final state goes to the other buffer: a neighbour tile may still
     * be reading its launch-start halo from state_a (nothing orders the
     * two in a one-step launch). The host swaps the buffers. */
#pragma unroll
    for (int k = 0; k < MK_SCPT; ++k) {
        if (act[k]) {
            const long long c = ix + (gy0 + w * MK_SCPT + k) * nx;
#pragma unroll
            for (int i = 0; i < MK_NS; ++i)
                a.state_b[(long long)i * a.npad + c] = y[k][i];
        }
    }
}

/* ------------------------------------------------------------------------
 * Persistent chains: disjoint chains of cells (cables built with
 * set_connections) laid end to end in the device's cell order, every
 * connection joining cells c and c + 1 (diff_arb_step, openclsim.cl:537-556,
 * for graphs of that shape).
 *
 * Each workgroup is one wave (64 threads) and owns ch.tile_len consecutive
 * cells (at most 64 * MK_CCPT; the last tile may be shorter) for every step
 * of the launch, thread t the cells t + 64 k, with their full state in
 * registers. One-wave tiles balance the work over the CUs to within one
 * wave. The tile's membrane potentials and one halo cell on each side live
 * in LDS (double-buffered by step). A cable longer than a tile spans
 * several: after each step a tile stores its first and last V as
 * step-tagged words (the exchange of mk_sheet); before the next step it
 * waits (bounded) for the value of each neighbour it is connected to. Tiles
 * whose boundary falls between two cables neither wait nor read there. As
 * in mk_sheet, part A of the cells' step runs before the halo wait and
 * part B (with the diffusion current) after it. Where the block driver
 * gives a whole cable to one workgroup, so that the longest cable sets the
 * launch's duration, here every cable is spread over as many tiles as its
 * length needs.
 *
 * A cell's diffusion current sums its (at most two) connection terms in the
 * cell's connection-list order, as mk_graph does.
 * ---------------------------------------------------------------------- */
#ifndef MK_CCPT
#define MK_CCPT 2
#endif
#ifndef MK_CWAVES
#define MK_CWAVES 4     /* waves per SIMD the register budget must allow */
#endif
#define MK_CT (64 * MK_CCPT)
typedef MkChainT<Real> MkChain;

/* Edge e (0 first cell, 1 last cell) of a tile, slot of step parity par. */
__device__ __forceinline__ MkXWord* mk_cedge(const MkChain& ch, int par,
                                             int tile, int e)
{
    return ch.edges + (((long long)par * ch.tiles + tile) * 2 + e) * MK_XW;
}

extern "C" __global__ void __launch_bounds__(64, MK_CWAVES)
mk_chain(const MkArgs a, const MkChain ch, const int s0, const int s1)
{
    __shared__ Real vc[2][MK_CT + 2];
    __shared__ int wg_flag, need[2];
    if (threadIdx.x == 0)
        wg_flag = mk_halted(a, a.steps[s1 - 1].row) ? 1 : 0;
    __syncthreads();
    if (wg_flag) return;

    const int tile = blockIdx.x;
    const long long c0 = (long long)tile * ch.tile_len;
    const int len = (int)(a.n - c0 < ch.tile_len ? a.n - c0 : ch.tile_len);
    const int t = threadIdx.x;

    Real y[MK_CCPT][MK_NS];
    Real fv[MK_CCPT][MK_NFA];
    Real gl[MK_CCPT], gr[MK_CCPT];
    unsigned cf[MK_CCPT];
    bool act[MK_CCPT], paced[MK_CCPT];
    Real q[MK_NQA];
#pragma unroll
    for (int k = 0; k < MK_CCPT; ++k) {
        const int li = t + 64 * k;
        const long long c = c0 + li;
        act[k] = li < len;
        const long long cc = act[k] ? c : 0;
        paced[k] = act[k] && mk_is_paced(a, cc, cc % a.nx, cc / a.nx);
#pragma unroll
        for (int i = 0; i < MK_NS; ++i)
            y[k][i] = act[k] ? a.state_a[(long long)i * a.npad + c] : (Real)0;
#pragma unroll
        for (int f = 0; f < MK_NF; ++f)
            fv[k][f] = act[k] ? a.fields[(long long)f * a.npad + c] : (Real)0;
        gl[k] = act[k] ? ch.gl[c] : (Real)0;
        gr[k] = act[k] ? ch.gr[c] : (Real)0;
        cf[k] = act[k] ? (unsigned)ch.cf[c] : 0u;
        if (act[k]) vc[0][li + 1] = y[k][MK_IVM];
        /* the neighbours this tile exchanges with: its first cell connected
         * to the previous tile's last, its last cell to the next tile's
         * first */
        if (li == 0) need[0] = tile > 0 && (cf[k] & 1u);
        if (li == len - 1) need[1] = tile + 1 < ch.tiles && (cf[k] & 2u);
    }
    __syncthreads();
    const bool nl = need[0] != 0, nr = need[1] != 0;
    /* halo at the start of the launch, from the global state */
    {
        const Real* V = a.state_a + (long long)MK_IVM * a.npad;
        if (t == 0) vc[0][0] = nl ? V[c0 - 1] : (Real)0;
        if (t == 1) vc[0][len + 1] = nr ? V[c0 + len] : (Real)0;
    }
    __syncthreads();

    int cur = 0;
    Real cy[MK_CCPT][MK_NCARRY];
    for (int s = s0; s < s1; ++s) {
        const int i = s - s0;
        const MkStep p = a.steps[s];
        const bool last = s + 1 == s1;
        const int par = (ch.g0 + i) & 1;
        /* Part A of every cell's step (all that does not depend on the
         * diffusion current) while the neighbours' edges are on their way */
#pragma unroll
        for (int k = 0; k < MK_CCPT; ++k) {
            if (act[k]) {
                const long long c = c0 + t + 64 * k;
                if (p.row >= 0 && a.nan_check && c == a.nan_cell &&
                    y[k][0] != y[k][0])
                    __atomic_store_n(a.halt, p.row, __ATOMIC_RELAXED);
                const Real pace = paced[k] ? p.pace : (Real)0;
                /* logged values are register copies, written out only at
                 * log steps */
                mk_cell_a<true>(y[k], p.t, p.dt, pace, fv[k], q, cy[k]);
                if (p.row >= 0) mk_write_log_part<0>(a, c, p.row, q);
            }
            /* one cell after the other: interleaving the cells' part A
             * buys no latency the other waves do not hide, and multiplies
             * the live registers */
            __builtin_amdgcn_sched_barrier(0);
        }
        if (i > 0 && (nl || nr)) {
            /* the connected neighbours' edge cells of the previous step
             * into the halo */
            const int step = ch.g0 + i;
            const int hpar = (step - 1) & 1;
            if ((t == 0 && nl) || (t == 1 && nr)) {
                Real v;
                if (mk_tget(t == 0 ? mk_cedge(ch, hpar, tile - 1, 1)
                                   : mk_cedge(ch, hpar, tile + 1, 0),
                            step, ch.spin_limit, v)) {
                    vc[cur][t == 0 ? 0 : len + 1] = v;
                } else {
                    __hip_atomic_store(ch.err, 1, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    wg_flag = 1;
                }
            }
            __syncthreads();
            if (wg_flag) return;
        }
        /* Part B: the diffusion current from the potentials at t, then the
         * rest of the step */
#pragma unroll
        for (int k = 0; k < MK_CCPT; ++k) {
            const int li = t + 64 * k;
            if (act[k]) {
                const long long c = c0 + li;
                const Real v = vc[cur][li + 1];
                const Real vl = vc[cur][li];
                const Real vr = vc[cur][li + 2];
                /* the connection terms in list order (mk_graph's sums) */
                Real id = (Real)0;
                if (cf[k] & 4u) {
                    if (cf[k] & 1u) id = id + gl[k] * (v - vl);
                    if (cf[k] & 2u) id = id + gr[k] * (v - vr);
                } else {
                    if (cf[k] & 2u) id = id + gr[k] * (v - vr);
                    if (cf[k] & 1u) id = id + gl[k] * (v - vl);
                }
                const Real pace = paced[k] ? p.pace : (Real)0;
                mk_cell_b<true>(y[k], p.t, p.dt, pace, id, fv[k], q, cy[k]);
                if (p.row >= 0) mk_write_log_part<1>(a, c, p.row, q);
                const Real vn = y[k][MK_IVM];
                vc[cur ^ 1][li + 1] = vn;
                if (!last) {
                    if (li == 0 && nl)
                        mk_tput(mk_cedge(ch, par, tile, 0), vn, ch.g0 + i + 1);
                    if (li == len - 1 && nr)
                        mk_tput(mk_cedge(ch, par, tile, 1), vn, ch.g0 + i + 1);
                }
            }
        }
        /* LDS writes of this step before the next step's reads (with a
         * neighbour to wait for, the barrier after the halo does it) */
        if (!(nl || nr)) __syncthreads();
        cur ^= 1;
    }
    /* the final state goes to the other buffer: a neighbour may still be
     * reading its launch-start halo from state_a. The host swaps them. */
#pragma unroll
    for (int k = 0; k < MK_CCPT; ++k) {
        if (act[k]) {
            const long long c = c0 + t + 64 * k;
#pragma unroll
            for (int i = 0; i < MK_NS; ++i)
                a.state_b[(long long)i * a.npad + c] = y[k][i];
        }
    }
}

/* ------------------------------------------------------------------------
 * Fiber-tissue junction (diff_step_fiber_tissue, openclsim.cl:601-628):
 * fiber row k's last cell (nfx - 1, k) connects to tissue cell (ctx,
 * cty + k) with conductance gft. The current gft (Vf - Vt) is written for
 * the fiber cell and its negative for the tissue cell, into arrays that
 * are zero elsewhere; the grid driver adds them after its stencil.
 * ---------------------------------------------------------------------- */
extern "C" __global__ void __launch_bounds__(64)
mk_junction(const Real* __restrict__ vf, const Real* __restrict__ vt,
            Real* __restrict__ jf, Real* __restrict__ jt, const long long nfx,
            const long long nfy, const long long ntx, const long long ctx,
            const long long cty, const Real gft)
{
    const long long k = (long long)blockIdx.x * 64 + threadIdx.x;
    if (k >= nfy) return;
    const long long iff = (nfx - 1) + k * nfx;
    const long long ift = ctx + (cty + k) * ntx;
    const Real i = gft * (vf[iff] - vt[ift]);
    jf[iff] = i;
    jt[ift] = -i;
}
