/*
 * mk_adjoint.hip -- discrete adjoint kernels (appended to a generated kernel
 * source together with mk_cell_vjp when gradients are requested).
 *
 * For the step y_{s+1} = F_s(y_s) of the whole tissue, s = 0..N-1, and the
 * loss
 *   L = sum_{s=0..N-1} sum_p (V(y_{s+1})[probe p] - target[s][p])^2,
 * the adjoint recursion is
 *   lam_N = dL/dy_N,   lam_s = (dF_s/dy_s)^T lam_{s+1} + dL/dy_s,
 * and a parameter gradient is sum_s (dF_s/dparam)^T lam_{s+1}.
 * (dF_s/dy_s)^T lam_{s+1} has two parts: the cell's own update, whose
 * transpose mk_cell_vjp applies (mu_s), and the coupling through the
 * diffusion current idiff_c = sum_j g_cj (V_c - V_j). With abar_s the
 * per-cell adjoint of idiff (also from mk_cell_vjp), the coupling part is
 * the same weighted graph Laplacian applied to abar_s, added to the V
 * component:  sum_j g_cj (abar_s[c] - abar_s[j]).
 *
 * One reverse step s assembles lam_{s+1} from mu_{s+1}, the Laplacian of
 * abar_{s+1} and the loss term at y_{s+1}, adds the conductance-gradient
 * terms of step s+1, then applies the cell VJP of step s at y_s, giving
 * mu_s and abar_s. lam itself is never stored. Two drivers:
 *
 *   mk_adj_step   one reverse step per launch over all cells (any size); the
 *                 Laplacian needs every cell's abar_{s+1}, hence the launch
 *                 boundary between steps;
 *   mk_adj_block  one workgroup per independent domain (cable, small sheet,
 *                 graph component), every reverse step of a segment in one
 *                 launch; V and abar of the domain live in LDS, one barrier
 *                 per step.
 *
 * Field gradients accumulate per cell in double precision. Conductance
 * gradients are summed over each workgroup (wave shuffles, then LDS) into
 * one double per workgroup, so every partial sum has a single writer and a
 * fixed order: results do not depend on scheduling.
 */
typedef MkAdjT<Real> MkAdj;

/* Sum over the workgroup of a double per thread; valid in thread 0. */
__device__ __forceinline__ double mk_block_sum(double v, double* lds)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < MK_BT / 64; ++k) r += lds[k];
    }
    return r;
}

/* Coupling transpose at cell c for abar_{s+1} (ab) and V(y_{s+1}) (V1):
 * returns sum_j g_cj (ab[c] - ab[j]) and adds the cell's terms of dL/dgx and
 * dL/dgy (graphs: dL/dscale, half from each end of a connection) to cx,
 * cy. Arrays are indexed as in the forward coupling: by cell on a grid or
 * a single stencil domain; graphs by CSR row `row` with neighbour indices
 * relative to the arrays, the cell itself at `self`. */
__device__ __forceinline__ Real mk_adj_couple(
    const MkArgs& a, const Real* __restrict__ V1, const Real* __restrict__ ab,
    int c, int row, int self, int ix, int iy, int nx, int ny, double& cx,
    double& cy)
{
    Real lv = (Real)0;
    (void)row; (void)self; (void)ix; (void)iy; (void)c; (void)nx; (void)ny;
#if MK_COUPLING == 1
    const Real ac = ab[c], v = V1[c];
    Real ax = (Real)0, ay = (Real)0, xs = (Real)0, ys = (Real)0;
    if (nx > 1) {
        if (ix > 0) { ax += ac - ab[c - 1]; xs += v - V1[c - 1]; }
        if (ix < nx - 1) { ax += ac - ab[c + 1]; xs += v - V1[c + 1]; }
    }
    if (ny > 1) {
        if (iy > 0) { ay += ac - ab[c - nx]; ys += v - V1[c - nx]; }
        if (iy < ny - 1) { ay += ac - ab[c + nx]; ys += v - V1[c + nx]; }
    }
    lv = a.gx * ax + a.gy * ay;
    cx += (double)ac * (double)xs;
    cy += (double)ac * (double)ys;
#elif MK_COUPLING == 2
    const Real ac = ab[c];
    if (nx > 1) {
        if (ix > 0) lv += a.gx_field[c - iy - 1] * (ac - ab[c - 1]);
        if (ix < nx - 1) lv += a.gx_field[c - iy] * (ac - ab[c + 1]);
    }
    if (ny > 1) {
        if (iy > 0) lv += a.gy_field[c - nx] * (ac - ab[c - nx]);
        if (iy < ny - 1) lv += a.gy_field[c] * (ac - ab[c + nx]);
    }
#elif MK_COUPLING == 3
    const Real ac = ab[self], v = V1[self];
    /* rows of one copy for a graph of interleaved copies (see mk_graph) */
    const int K = a.gcopies;
    int lane = 0;
    if (K > 1) {
        row = self / K;
        lane = self - row * K;
    }
    int e = a.csr_ptr[row];
    const int e1 = a.csr_ptr[row + 1];
    /* four connections per round, loads first, sums in list order (as in
     * mk_graph) */
    for (; e + 4 <= e1; e += 4) {
        int nb[4];
        Real gg[4], an[4], vn[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int t = a.csr_idx[e + k];
            nb[k] = (t < 0 ? ~t : t) * K + lane;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            gg[k] = a.csr_g[e + k];
            an[k] = ab[nb[k]];
            vn[k] = V1[nb[k]];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            lv += gg[k] * (ac - an[k]);
            cx += 0.5 * (double)gg[k] * ((double)ac - (double)an[k])
                  * ((double)v - (double)vn[k]);
        }
    }
    for (; e < e1; ++e) {
        int nb = a.csr_idx[e];
        if (nb < 0) nb = ~nb;
        nb = nb * K + lane;
        const Real g = a.csr_g[e];
        lv += g * (ac - ab[nb]);
        cx += 0.5 * (double)g * ((double)ac - (double)ab[nb])
              * ((double)v - (double)V1[nb]);
    }
#endif
    return lv;
}

#if MK_COUPLING == 3
/* mk_adj_couple and mk_graph in one walk of graph row c (the one-step
 * kernel: row = self = c): returns the coupling transpose at abar_{s+1}
 * (ab) and V(y_{s+1}) (V1), adds the cell's dL/dscale terms to cx, and
 * sets idiff to the forward current at V(y_s) (V0). Each sum takes the
 * same additions in the same order as in the separate walks. */
__device__ __forceinline__ Real mk_adj_couple_graph(
    const MkArgs& a, const Real* __restrict__ V1, const Real* __restrict__ ab,
    const Real* __restrict__ V0, int c, double& cx, Real& idiff)
{
    const Real ac = ab[c], v = V1[c], v0 = V0[c];
    Real lv = (Real)0, id = (Real)0;
    /* rows of one copy for a graph of interleaved copies (see mk_graph) */
    const int K = a.gcopies;
    const int row = K > 1 ? c / K : c, lane = c - row * K;
    int e = a.csr_ptr[row];
    const int e1 = a.csr_ptr[row + 1];
    for (; e + 4 <= e1; e += 4) {
        int t[4];
        Real gg[4], an[4], vn[4], wn[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) t[k] = a.csr_idx[e + k];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int nb = (t[k] < 0 ? ~t[k] : t[k]) * K + lane;
            gg[k] = a.csr_g[e + k];
            an[k] = ab[nb];
            vn[k] = V1[nb];
            wn[k] = V0[nb];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            lv += gg[k] * (ac - an[k]);
            cx += 0.5 * (double)gg[k] * ((double)ac - (double)an[k])
                  * ((double)v - (double)vn[k]);
            if (t[k] >= 0) id = id + gg[k] * (v0 - wn[k]);
            else id = id - gg[k] * (wn[k] - v0);
        }
    }
    for (; e < e1; ++e) {
        const int t = a.csr_idx[e];
        const int nb = (t < 0 ? ~t : t) * K + lane;
        const Real g = a.csr_g[e];
        lv += g * (ac - ab[nb]);
        cx += 0.5 * (double)g * ((double)ac - (double)ab[nb])
              * ((double)v - (double)V1[nb]);
        if (t >= 0) id = id + g * (v0 - V0[nb]);
        else id = id - g * (V0[nb] - v0);
    }
    idiff = id;
    return lv;
}
#endif

/* Loss term dL/dV at cell c for V(y_{s+1}) = v, target row `row`. */
__device__ __forceinline__ Real mk_adj_loss(const MkAdj& j, int c, Real v,
                                            int row)
{
    if (row < 0) return (Real)0;
    const int q = j.probe_of[c];
    if (q < 0) return (Real)0;
    const double d = (double)v - j.target[(long long)row * j.n_probe + q];
    j.loss_acc[q] += d * d;
    return (Real)(2.0 * d);
}

/* Cell VJP of step p at state y (own cell), V the potentials of y_s indexed
 * as in the forward coupling. */
__device__ __forceinline__ void mk_adj_cell(
    const MkArgs& a, const MkStep& p, const Real* y, const Real* fv,
    const Real* __restrict__ V, int c, int row, int self, int ix, int iy,
    int nx, int ny, const Real* lam, Real* mu, Real& abo, Real* fb)
{
    (void)row; (void)self; (void)nx; (void)ny;
#if MK_COUPLING == 1
    const Real idiff = mk_stencil(V, c, ix, iy, nx, ny, a.gx, a.gy);
#elif MK_COUPLING == 2
    const Real idiff = mk_hetero(V, a.gx_field, a.gy_field, c, ix, iy, nx,
                                 ny);
#elif MK_COUPLING == 3
    const Real idiff = mk_graph(V, a.csr_ptr, a.csr_idx, a.csr_g, row, self,
                                a.gcopies);
#else
    const Real idiff = (Real)0;
    (void)V;
#endif
    const Real pace = mk_is_paced(a, c, ix, iy) ? p.pace : (Real)0;
    mk_cell_vjp(y, p.t, p.dt, pace, idiff, fv, lam, mu, &abo, fb);
}

/* ---------------------------------------------------------------------- */
extern "C" __global__ void __launch_bounds__(MK_BT, MK_ADJ_WAVES)
mk_adj_step(const MkArgs a, const MkAdj j, const int s)
{
    __shared__ double red[2][MK_BT / 64];
    const int c = blockIdx.x * MK_BT + threadIdx.x;
    double cx = 0.0, cy = 0.0;
    if (c < (int)a.n) {
        const int nx = (int)a.nx, ny = (int)a.ny;
        const int iy = (int)((unsigned)c / (unsigned)nx), ix = c - iy * nx;
        const long long np = a.npad;
        Real lam[MK_NS];
        if (j.mu_in) {
            const Real* __restrict__ m = j.mu_in + c;
#pragma unroll
            for (int i = 0; i < MK_NS; ++i) lam[i] = m[i * np];
        } else {
#pragma unroll
            for (int i = 0; i < MK_NS; ++i) lam[i] = (Real)0;
        }
        const Real* __restrict__ V1 = j.yn1 + MK_IVM * np;
        Real lv = mk_adj_loss(j, c, V1[c], j.target_row);
#if MK_COUPLING == 3
        /* both walks of the cell's connections in one */
        Real idiff0 = (Real)0;
        const bool fused = j.ab_in && j.yn;
        if (fused)
            lv += mk_adj_couple_graph(a, V1, j.ab_in, j.yn + MK_IVM * np, c,
                                      cx, idiff0);
        else if (j.ab_in)
#else
        if (j.ab_in)
#endif
            lv += mk_adj_couple(a, V1, j.ab_in, c, c, c, ix, iy, nx, ny, cx,
                                cy);
        lam[MK_IVM] += lv;
        if (j.yn) {
            Real y[MK_NS], mu[MK_NS], fv[MK_NFA], fb[MK_NFA], abo;
            const Real* __restrict__ yp = j.yn + c;
#pragma unroll
            for (int i = 0; i < MK_NS; ++i) y[i] = yp[i * np];
#pragma unroll
            for (int k = 0; k < MK_NF; ++k) fv[k] = a.fields[k * np + c];
#if MK_COUPLING == 3
            if (!fused)
                idiff0 = mk_graph(j.yn + MK_IVM * np, a.csr_ptr, a.csr_idx,
                                  a.csr_g, c, c, a.gcopies);
            {
                const MkStep& p = a.steps[s];
                const Real pace = mk_is_paced(a, c, ix, iy) ? p.pace
                                                            : (Real)0;
                mk_cell_vjp(y, p.t, p.dt, pace, idiff0, fv, lam, mu, &abo,
                            fb);
            }
#else
            mk_adj_cell(a, a.steps[s], y, fv, j.yn + MK_IVM * np, c, c, c, ix,
                        iy, nx, ny, lam, mu, abo, fb);
#endif
            Real* __restrict__ mo = j.mu_out + c;
#pragma unroll
            for (int i = 0; i < MK_NS; ++i) mo[i * np] = mu[i];
            j.ab_out[c] = abo;
#pragma unroll
            for (int k = 0; k < MK_NF; ++k) j.gfield[k * np + c] += (double)fb[k];
        }
    }
#if MK_COUPLING == 1 || MK_COUPLING == 3
    if (j.ab_in) {
        const double sx = mk_block_sum(cx, red[0]);
        const double sy = mk_block_sum(cy, red[1]);
        if (threadIdx.x == 0) {
            j.gx_acc[blockIdx.x] += sx;
            j.gy_acc[blockIdx.x] += sy;
        }
    }
#else
    (void)red;
#endif
}

/* Every reverse step of one segment [s0, s1) for one domain per workgroup.
 * j.yn is the segment's trajectory: y_{s0 + k} at j.yn + k * n_state * npad
 * for k = 0..s1-s0. j.mu_out holds mu of each cell (read and written in
 * place: a cell only reads its own); j.ab_out carries abar of each cell
 * from one launch to the next. flags: bit 0 = this segment ends the run (no
 * mu or abar of step N yet), bit 1 = this segment starts the run (add the
 * conductance terms of step 0 at the end). */
extern "C" __global__ void __launch_bounds__(MK_BT, MK_ADJ_WAVES)
mk_adj_block(const MkArgs a, const MkAdj j, const int s0, const int s1,
             const int flags)
{
    __shared__ Real vsh[2][MK_BCELLS];
    __shared__ Real ash[2][MK_BCELLS];
    __shared__ double red[2][MK_BT / 64];
    const long long off = a.dom_off[blockIdx.x];
    const int len = (int)(a.dom_off[blockIdx.x + 1] - off);
    const long long SN = (long long)MK_NS * a.npad;
    const bool last_seg = flags & 1, first_seg = flags & 2;
    double cx = 0.0, cy = 0.0;
    /* V(y_{s1}) and abar_{s1} (zero at the end of the run) */
    int cur = 0;
    const int nx = (int)a.nx, ny = (int)a.ny;
    for (int li = threadIdx.x; li < len; li += MK_BT) {
        const int c = a.dom_cells ? a.dom_cells[off + li] : (int)(off + li);
        vsh[0][li] = j.yn[(long long)(s1 - s0) * SN
                          + (long long)MK_IVM * a.npad + c];
        ash[0][li] = last_seg ? (Real)0 : j.ab_out[c];
    }
    __syncthreads();
    for (int s = s1 - 1; s >= s0; --s) {
        const Real* __restrict__ ys = j.yn + (long long)(s - s0) * SN;
        /* V(y_s) of the domain into the other buffer */
        for (int li = threadIdx.x; li < len; li += MK_BT) {
            const int c = a.dom_cells ? a.dom_cells[off + li]
                                      : (int)(off + li);
            vsh[cur ^ 1][li] = ys[(long long)MK_IVM * a.npad + c];
        }
        __syncthreads();
        const MkStep p = a.steps[s];
        const bool fresh = last_seg && s == s1 - 1;
        for (int li = threadIdx.x; li < len; li += MK_BT) {
            const int c = a.dom_cells ? a.dom_cells[off + li]
                                      : (int)(off + li);
            const int iy = (int)((unsigned)c / (unsigned)nx);
            const int ix = c - iy * nx;
            const int row = (int)(off + li);
            Real lam[MK_NS], y[MK_NS], mu[MK_NS], fv[MK_NFA], fb[MK_NFA], abo;
#pragma unroll
            for (int i = 0; i < MK_NS; ++i)
                lam[i] = fresh ? (Real)0 : j.mu_out[(long long)i * a.npad + c];
            Real lv = mk_adj_loss(j, c, vsh[cur][li], s);
            if (!fresh)
                lv += mk_adj_couple(a, vsh[cur], ash[cur], li, row, li, ix,
                                    iy, nx, ny, cx, cy);
            lam[MK_IVM] += lv;
#pragma unroll
            for (int i = 0; i < MK_NS; ++i)
                y[i] = ys[(long long)i * a.npad + c];
#pragma unroll
            for (int k = 0; k < MK_NF; ++k)
                fv[k] = a.fields[(long long)k * a.npad + c];
            mk_adj_cell(a, p, y, fv, vsh[cur ^ 1], li, row, li, ix, iy, nx,
                        ny, lam, mu, abo, fb);
#pragma unroll
            for (int i = 0; i < MK_NS; ++i)
                j.mu_out[(long long)i * a.npad + c] = mu[i];
            ash[cur ^ 1][li] = abo;
#pragma unroll
            for (int k = 0; k < MK_NF; ++k)
                j.gfield[(long long)k * a.npad + c] += (double)fb[k];
        }
        cur ^= 1;
        __syncthreads();
    }
    /* Carry abar_{s0} to the previous segment; at the start of the run add
     * the conductance terms of step 0 instead. */
    for (int li = threadIdx.x; li < len; li += MK_BT) {
        const int c = a.dom_cells ? a.dom_cells[off + li] : (int)(off + li);
        if (first_seg) {
            const int iy = (int)((unsigned)c / (unsigned)nx);
            (void)mk_adj_couple(a, vsh[cur], ash[cur], li, (int)(off + li),
                                li, c - iy * nx, iy, nx, ny, cx, cy);
        } else {
            j.ab_out[c] = ash[cur][li];
        }
    }
#if MK_COUPLING == 1 || MK_COUPLING == 3
    const double sx = mk_block_sum(cx, red[0]);
    const double sy = mk_block_sum(cy, red[1]);
    if (threadIdx.x == 0) {
        j.gx_acc[blockIdx.x] += sx;
        j.gy_acc[blockIdx.x] += sy;
    }
#else
    (void)red; (void)cx; (void)cy;
#endif
}

/* V at the probe cells of state yn (trace row s). */
extern "C" __global__ void __launch_bounds__(64)
mk_adj_probe(const MkArgs a, const MkAdj j, const int s)
{
    const int q = blockIdx.x * 64 + threadIdx.x;
    if (q >= j.n_probe) return;
    j.trace[(long long)s * j.n_probe + q] =
        (double)j.yn[(long long)MK_IVM * a.npad + j.probes[q]];
}

/* The field gradients in the caller's cell order, for the download:
 * dst[k * n + c] = src[k * npad + order[c]], k < nf, c < n (the host checks
 * 0 <= order[c] < n). */
extern "C" __global__ void __launch_bounds__(256)
mk_adj_gather(const double* __restrict__ src,
              const long long* __restrict__ order, double* __restrict__ dst,
              const long long n, const long long npad, const int nf)
{
    const long long total = n * nf;
    const long long stride = (long long)gridDim.x * 256;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total;
         i += stride) {
        const long long k = i / n;
        dst[i] = src[k * npad + order[i - k * n]];
    }
}

#if MK_ADJ_SHEET
/* ------------------------------------------------------------------------
 * Persistent reverse sweep on a 2D stencil sheet: every reverse step of a
 * segment [s0, s1) in one cooperative launch.
 *
 * Each workgroup owns a tile of 64 x MK_ATH cells (MK_ATH = 4 * MK_ACPT);
 * thread (lx, w) handles cells (x0 + lx, y0 + w * MK_ACPT + k), k <
 * MK_ACPT, one after the other. MK_ADJ_SWAVES workgroups per CU fit into
 * LDS, and the registers are budgeted for as many waves per SIMD. mu of the tile stays in LDS for the whole
 * segment, abar of the tile and a one-cell ring around it too. The
 * potentials V_s and V_{s+1} come from the trajectory, complete in global
 * memory, so only abar crosses tiles: after each step a tile stores the
 * abar of its four edges as step-tagged words (the exchange of mk_sheet),
 * and before the next step each thread of its ring waits for its
 * neighbour's value of that step.
 *
 * A step has two phases. Phase 1 computes, for every cell of the tile, the
 * coupling transpose gx (abar_c - abar_x) + gy (...) from abar_{s+1} and
 * the cell's terms of dL/dgx, dL/dgy. After a barrier, phase 2 assembles
 * lam = mu_{s+1} + coupling + loss, applies the cell VJP at y_s and writes
 * mu_s and abar_s in place. The barrier orders every read of abar_{s+1}
 * before the first write of abar_s, so one abar buffer suffices.
 *
 * j.yn is the segment's trajectory (y_{s0 + k} at j.yn + k * n_state *
 * npad); j.mu_out and j.ab_out carry mu and abar between segments. flags:
 * bit 0 = this segment ends the run (mu, abar of step s1 are zero), bit 1 =
 * it starts the run (add the conductance terms of step 0 at the end). Every
 * wait is bounded (sh.spin_limit); on a timeout the tile sets sh.err and
 * returns, and its neighbours then time out too.
 * ---------------------------------------------------------------------- */
#define MK_ATH (4 * MK_ACPT)

/* Coupling transpose at the tile cell (r, lx) from abar in LDS (ring
 * included) and V_{s+1} in global memory; adds the cell's conductance
 * terms to cx, cy. Same arithmetic as mk_adj_couple on the grid. */
__device__ __forceinline__ Real mk_adj_sheet_couple(
    const MkArgs& a, const Real (*at)[66], const Real* __restrict__ V1,
    int c, int r, int lx, int ix, int iy, int nx, int ny, double& cx,
    double& cy)
{
    const Real ac = at[r + 1][lx + 1], v = V1[c];
    Real ax = (Real)0, ay = (Real)0, xs = (Real)0, ys = (Real)0;
    if (nx > 1) {
        if (ix > 0) { ax += ac - at[r + 1][lx]; xs += v - V1[c - 1]; }
        if (ix < nx - 1) { ax += ac - at[r + 1][lx + 2]; xs += v - V1[c + 1]; }
    }
    if (ny > 1) {
        if (iy > 0) { ay += ac - at[r][lx + 1]; ys += v - V1[c - nx]; }
        if (iy < ny - 1) { ay += ac - at[r + 2][lx + 1]; ys += v - V1[c + nx]; }
    }
    cx += (double)ac * (double)xs;
    cy += (double)ac * (double)ys;
    return a.gx * ax + a.gy * ay;
}

/* Value k of abar edge e (0 bottom row, 1 top row, 2 left column, 3 right
 * column) of a tile, slot of step parity par (the layout of mk_edge). */
__device__ __forceinline__ MkXWord* mk_adj_edge(const MkSheet& sh, int par,
                                                int tile, int e, int k)
{
    return sh.edges + (((long long)(par * sh.tiles_x * sh.tiles_y + tile) * 4
                        + e) * 64 + k) * MK_XW;
}

/* Waits until each neighbour in nb[4] has stored the abar edges of local
 * step `step` (tagged step; its edge facing this tile, value 0). Returns
 * false on a timeout. */
__device__ __forceinline__ bool mk_adj_sheet_wait(const MkSheet& sh,
                                                  const int nb[4], int step,
                                                  int* flag)
{
    if (threadIdx.x < 4) {
        /* the edge of neighbour t facing this tile is t ^ 1 */
        const int nbt = nb[threadIdx.x];
        Real v;
        if (nbt >= 0 &&
            !mk_tget(mk_adj_edge(sh, (step - 1) & 1, nbt, threadIdx.x ^ 1, 0),
                     step, sh.spin_limit, v)) {
            __hip_atomic_store(sh.err, 1, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            *flag = 1;
        }
    }
    __syncthreads();
    return !*flag;
}

/* The neighbours' abar edges of local step `step` (tagged step) into the
 * ring of at, each ring thread waiting for its own value. Returns false on
 * a timeout. */
__device__ __forceinline__ bool mk_adj_sheet_halo(
    const MkSheet& sh, const int nb[4], int step, Real (*at)[66], int* flag,
    int gx0, int gy0, int nx, int ny)
{
    const int par = (step - 1) & 1;
    const int t = threadIdx.x;
    int nbt = -1, e = 0, k = 0, lr = 0, lc = 0;
    if (t < 64) {
        nbt = nb[0]; e = 1; k = t; lr = 0; lc = t + 1;
    } else if (t < 128) {
        nbt = nb[1]; e = 0; k = t - 64; lr = MK_ATH + 1; lc = t - 63;
    } else if (t < 128 + MK_ATH) {
        nbt = nb[2]; e = 3; k = t - 128; lr = t - 127; lc = 0;
    } else if (t < 128 + 2 * MK_ATH) {
        nbt = nb[3]; e = 2; k = t - 128 - MK_ATH; lr = t - 127 - MK_ATH;
        lc = 65;
    }
    /* only cells of the sheet are written (not the columns of a partial
     * last tile column or the rows of a partial last tile row) */
    if (nbt >= 0 && (e < 2 ? gx0 + k < nx : gy0 + k < ny)) {
        Real v;
        if (mk_tget(mk_adj_edge(sh, par, nbt, e, k), step, sh.spin_limit, v)) {
            at[lr][lc] = v;
        } else {
            __hip_atomic_store(sh.err, 1, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            *flag = 1;
        }
    }
    __syncthreads();
    return !*flag;
}
extern "C" __global__ void __launch_bounds__(256, MK_ADJ_SWAVES)
mk_adj_sheet(const MkArgs a, const MkAdj j, const MkSheet sh, const int s0,
             const int s1, const int flags)
{
    __shared__ Real mul[MK_NS][MK_ATH][64];
    __shared__ Real at[MK_ATH + 2][66];
    __shared__ double red[2][4];
    __shared__ int wg_flag;
    if (threadIdx.x == 0) wg_flag = 0;
    const bool last_seg = flags & 1, first_seg = flags & 2;
    const int tile = blockIdx.x;
    const int tx = tile % sh.tiles_x, ty = tile / sh.tiles_x;
    const int lx = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nx = (int)a.nx, ny = (int)a.ny;
    const int ix = tx * 64 + lx, gy0 = ty * MK_ATH;
    const long long np = a.npad;
    const long long SN = (long long)MK_NS * np;
    const int nb[4] = {ty > 0 ? tile - sh.tiles_x : -1,
                       ty + 1 < sh.tiles_y ? tile + sh.tiles_x : -1,
                       tx > 0 ? tile - 1 : -1,
                       tx + 1 < sh.tiles_x ? tile + 1 : -1};
    double cx = 0.0, cy = 0.0;

    /* mu and abar of step s1, with abar's ring from global memory */
    for (int k = 0; k < MK_ACPT; ++k) {
        const int r = w * MK_ACPT + k, iy = gy0 + r;
        const bool act = ix < nx && iy < ny;
        const int c = act ? ix + iy * nx : 0;
        for (int i = 0; i < MK_NS; ++i)
            mul[i][r][lx] = (act && !last_seg) ? j.mu_out[i * np + c] : (Real)0;
        at[r + 1][lx + 1] = (act && !last_seg) ? j.ab_out[c] : (Real)0;
    }
    {
        const int t = threadIdx.x;
        int hx = -1, hy = -1, lr = 0, lc = 0;
        if (t < 66) {
            hx = tx * 64 - 1 + t; hy = gy0 - 1; lr = 0; lc = t;
        } else if (t < 132) {
            hx = tx * 64 - 1 + (t - 66); hy = gy0 + MK_ATH; lr = MK_ATH + 1;
            lc = t - 66;
        } else if (t < 132 + MK_ATH) {
            hx = tx * 64 - 1; hy = gy0 + (t - 132); lr = t - 132 + 1; lc = 0;
        } else if (t < 132 + 2 * MK_ATH) {
            hx = tx * 64 + 64; hy = gy0 + (t - 132 - MK_ATH);
            lr = t - 132 - MK_ATH + 1; lc = 65;
        }
        if (t < 132 + 2 * MK_ATH)
            at[lr][lc] = (!last_seg && hx >= 0 && hy >= 0 && hx < nx && hy < ny)
                             ? j.ab_out[hx + hy * nx] : (Real)0;
    }
    __syncthreads();

    const int nsteps = s1 - s0;
    for (int i = 0; i < nsteps; ++i) {
        const int s = s1 - 1 - i;
        if (i > 0 && !mk_adj_sheet_halo(sh, nb, i, at, &wg_flag, tx * 64, gy0,
                                        nx, ny))
            return;
        const Real* __restrict__ ys = j.yn + (long long)(s - s0) * SN;
        const Real* __restrict__ V0 = ys + MK_IVM * np;       /* V_s */
        const Real* __restrict__ V1 = V0 + SN;                 /* V_{s+1} */
        /* Phase 1: coupling transpose from abar_{s+1} */
        Real lc[MK_ACPT];
#pragma unroll
        for (int k = 0; k < MK_ACPT; ++k) {
            const int r = w * MK_ACPT + k, iy = gy0 + r;
            lc[k] = (Real)0;
            if (ix < nx && iy < ny)
                lc[k] = mk_adj_sheet_couple(a, at, V1, ix + iy * nx, r, lx, ix,
                                            iy, nx, ny, cx, cy);
        }
        __syncthreads();
        /* Phase 2: lam, the cell VJP at y_s, mu_s and abar_s in place */
        const MkStep p = a.steps[s];
        const int par = i & 1;
#pragma unroll 1
        for (int k = 0; k < MK_ACPT; ++k) {
            const int r = w * MK_ACPT + k, iy = gy0 + r;
            if (!(ix < nx && iy < ny)) continue;
            const int c = ix + iy * nx;
            Real lam[MK_NS], y[MK_NS], mu[MK_NS], fv[MK_NFA], fb[MK_NFA], abo;
#pragma unroll
            for (int q = 0; q < MK_NS; ++q) {
                lam[q] = mul[q][r][lx];
                y[q] = ys[q * np + c];
            }
#pragma unroll
            for (int f = 0; f < MK_NF; ++f) fv[f] = a.fields[f * np + c];
            Real lv = mk_adj_loss(j, c, V1[c], s);
            Real ck = (Real)0;
#pragma unroll
            for (int q = 0; q < MK_ACPT; ++q) if (q == k) ck = lc[q];
            lv += ck;
            lam[MK_IVM] += lv;
            mk_adj_cell(a, p, y, fv, V0, c, c, c, ix, iy, nx, ny, lam, mu, abo,
                        fb);
#pragma unroll
            for (int q = 0; q < MK_NS; ++q) mul[q][r][lx] = mu[q];
            at[r + 1][lx + 1] = abo;
            if (r == 0) mk_tput(mk_adj_edge(sh, par, tile, 0, lx), abo, i + 1);
            if (r == MK_ATH - 1)
                mk_tput(mk_adj_edge(sh, par, tile, 1, lx), abo, i + 1);
            if (lx == 0) mk_tput(mk_adj_edge(sh, par, tile, 2, r), abo, i + 1);
            if (lx == 63) mk_tput(mk_adj_edge(sh, par, tile, 3, r), abo, i + 1);
#pragma unroll
            for (int f = 0; f < MK_NF; ++f) j.gfield[f * np + c] += (double)fb[f];
        }
        /* the next step's barrier (after its halo) orders this step's
         * LDS writes before its reads */
    }
    /* mu and abar of step s0 for the previous segment; at the start of the
     * run, the conductance terms of step 0 instead */
    if (first_seg) {
        if (!mk_adj_sheet_halo(sh, nb, nsteps, at, &wg_flag, tx * 64, gy0, nx,
                               ny))
            return;
        const Real* __restrict__ V0 = j.yn + MK_IVM * np;
        for (int k = 0; k < MK_ACPT; ++k) {
            const int r = w * MK_ACPT + k, iy = gy0 + r;
            if (ix < nx && iy < ny)
                (void)mk_adj_sheet_couple(a, at, V0, ix + iy * nx, r, lx, ix,
                                          iy, nx, ny, cx, cy);
        }
    } else {
        /* A neighbour reads its launch-start ring from j.ab_out; it has
         * done so once it has finished its first step (with one step per
         * segment nothing else orders that read before these writes). */
        if (nsteps == 1 && !mk_adj_sheet_wait(sh, nb, 1, &wg_flag)) return;
        for (int k = 0; k < MK_ACPT; ++k) {
            const int r = w * MK_ACPT + k, iy = gy0 + r;
            if (!(ix < nx && iy < ny)) continue;
            const int c = ix + iy * nx;
            for (int q = 0; q < MK_NS; ++q) j.mu_out[q * np + c] = mul[q][r][lx];
            j.ab_out[c] = at[r + 1][lx + 1];
        }
    }
    /* Conductance partial sums: one per tile */
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        cx += __shfl_down(cx, o, 64);
        cy += __shfl_down(cy, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = cx;
        red[1][threadIdx.x >> 6] = cy;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        j.gx_acc[tile] += red[0][0] + red[0][1] + red[0][2] + red[0][3];
        j.gy_acc[tile] += red[1][0] + red[1][1] + red[1][2] + red[1][3];
    }
}
#endif
