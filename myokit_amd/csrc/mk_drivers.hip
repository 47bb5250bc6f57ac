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
 * SoA state, and three drivers cover the geometries:
 *
 *   mk_cells  uncoupled cells: one thread per cell keeps its state in
 *             registers across every step of the launch (no synchronisation);
 *   mk_block  one workgroup per independent domain (a cable, a small sheet,
 *             a connected component of a graph): V is exchanged through LDS
 *             with one barrier per step, all steps of the launch;
 *   mk_grid   any size: one fused step per launch, V of neighbours read from
 *             the previous state buffer, new state written to the other
 *             buffer (ping-pong).
 *
 * Halting: at a log step, the thread of cell 0 records the log row in
 * halt[0] if state 0 is NaN (openclsim.c:1087-1089). The host ends launches
 * at log steps, and a launch returns at once if halt[0] holds a row other
 * than its own, so steps after the halting step are never taken.
 */
typedef MkArgsT<Real> MkArgs;
typedef MkStepT<Real> MkStep;

#define MK_NFA (MK_NF > 0 ? MK_NF : 1)
#define MK_NQA (MK_NQ > 0 ? MK_NQ : 1)

__device__ __forceinline__ bool mk_is_paced(const MkArgs& a, long long c,
                                            long long ix, long long iy)
{
    if (a.pace_mode == 0) return true;
    if (a.pace_mode == 1)
        return ix >= a.px0 && ix < a.px1 && iy >= a.py0 && iy < a.py1;
    if (a.pace_mode == 2) return a.pace_mask[c] != 0;
    return false;
}

/* Homogeneous zero-flux stencil (diff_step, openclsim.cl:384-435); V is
 * indexed by cell id, same operation order as the reference. */
__device__ __forceinline__ Real mk_stencil(
    const Real* __restrict__ V, long long c, long long ix, long long iy,
    long long nx, long long ny, Real gx, Real gy)
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
__device__ __forceinline__ Real mk_hetero(
    const Real* __restrict__ V, const Real* __restrict__ gxf,
    const Real* __restrict__ gyf, long long c, long long ix, long long iy,
    long long nx, long long ny)
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
 * atomics happen to run in list order. V is indexed by `base + j`. */
__device__ __forceinline__ Real mk_graph(
    const Real* __restrict__ V, const int* __restrict__ ptr,
    const int* __restrict__ idx, const Real* __restrict__ g,
    long long row, long long self)
{
    const Real v = V[self];
    Real i = (Real)0;
    const int e1 = ptr[row + 1];
    for (int e = ptr[row]; e < e1; ++e) {
        const int j = idx[e];
        if (j >= 0) i = i + g[e] * (v - V[j]);
        else i = i - g[e] * (V[~j] - v);
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

__device__ __forceinline__ bool mk_halted(const MkArgs& a, int own_row)
{
    const int h = __atomic_load_n(a.halt, __ATOMIC_RELAXED);
    return h >= 0 && h != own_row;
}

/* ------------------------------------------------------------------------
 * Uncoupled cells: one thread per cell, state in registers for all steps.
 * ---------------------------------------------------------------------- */
extern "C" __global__ void __launch_bounds__(MK_BT)
mk_cells(const MkArgs a, const int s0, const int s1)
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
        if (p.row >= 0) {
            if (a.nan_check && c == 0 && y[0] != y[0])
                __atomic_store_n(a.halt, p.row, __ATOMIC_RELAXED);
            mk_cell<true>(y, p.t, p.dt, pace, (Real)0, fv, q);
            mk_write_log(a, c, p.row, q);
        } else {
            mk_cell<false>(y, p.t, p.dt, pace, (Real)0, fv, q);
        }
    }
#pragma unroll
    for (int i = 0; i < MK_NS; ++i) a.state_a[(long long)i * a.npad + c] = y[i];
}

/* ------------------------------------------------------------------------
 * One workgroup per domain; V exchanged through LDS, one barrier per step.
 * ---------------------------------------------------------------------- */
extern "C" __global__ void __launch_bounds__(MK_BT)
mk_block(const MkArgs a, const int s0, const int s1)
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
            const Real idiff = mk_stencil(vsh[cur], li, ix, iy, nx, ny,
                                          a.gx, a.gy);
#elif MK_COUPLING == 2
            const Real idiff = mk_hetero(vsh[cur], a.gx_field, a.gy_field,
                                         li, ix, iy, nx, ny);
#elif MK_COUPLING == 3
            const Real idiff = mk_graph(vsh[cur], a.csr_ptr, a.csr_idx,
                                        a.csr_g, off + li, li);
#else
            const Real idiff = (Real)0;
#endif
            const Real pace = mk_is_paced(a, c, ix, iy) ? p.pace : (Real)0;
            if (p.row >= 0) {
                if (a.nan_check && c == 0 && y[0] != y[0])
                    __atomic_store_n(a.halt, p.row, __ATOMIC_RELAXED);
                mk_cell<true>(y, p.t, p.dt, pace, idiff, fv, q);
                mk_write_log(a, c, p.row, q);
            } else {
                mk_cell<false>(y, p.t, p.dt, pace, idiff, fv, q);
            }
#pragma unroll
            for (int i = 0; i < MK_NS; ++i)
                a.state_a[(long long)i * a.npad + c] = y[i];
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

extern "C" __global__ void __launch_bounds__(MK_BT)
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
    const Real idiff = mk_stencil(V, c, ix, iy, a.nx, a.ny, a.gx, a.gy);
#elif MK_COUPLING == 2
    const Real idiff = mk_hetero(V, a.gx_field, a.gy_field, c, ix, iy,
                                 a.nx, a.ny);
#elif MK_COUPLING == 3
    const Real idiff = mk_graph(V, a.csr_ptr, a.csr_idx, a.csr_g, c, c);
#else
    const Real idiff = (Real)0;
#endif
    const Real pace = mk_is_paced(a, c, ix, iy) ? p.pace : (Real)0;
    if (p.row >= 0) {
        if (a.nan_check && c == 0 && y[0] != y[0])
            __atomic_store_n(a.halt, p.row, __ATOMIC_RELAXED);
        mk_cell<true>(y, p.t, p.dt, pace, idiff, fv, q);
        mk_write_log(a, c, p.row, q);
    } else {
        mk_cell<false>(y, p.t, p.dt, pace, idiff, fv, q);
    }
#pragma unroll
    for (int i = 0; i < MK_NS; ++i) dst[(long long)i * a.npad + c] = y[i];
}
