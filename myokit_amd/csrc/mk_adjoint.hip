/*
 * mk_adjoint.hip -- discrete adjoint kernels (appended to a generated kernel
 * source together with mk_cell_vjp when gradients are requested).
 *
 * For the step y_{s+1} = F(y_s) of the whole tissue, with the loss
 *   L = sum_s sum_p (V_s[probe p] - target[s-1][p])^2,   s = 1..N,
 * the adjoint recursion is lam_N = dL/dy_N and
 *   lam_s = (dF/dy_s)^T lam_{s+1} + dL/dy_s.
 * dF/dy_s has two parts: each cell's own update (mk_cell_vjp) and the
 * coupling through the diffusion current idiff_c = sum_j g_cj (V_c - V_j).
 * The transpose of the coupling applied to the per-cell adjoints abar of
 * idiff is the same weighted graph Laplacian, sum_j g_cj (abar_c - abar_j),
 * added to the V component. mk_adj_vjp computes the per-cell part;
 * mk_adj_couple adds the coupling, the loss gradient and the conductance
 * gradients. Gradient and loss accumulators are double precision in every
 * precision, and each cell accumulates its own entries (no atomics, so the
 * result does not depend on scheduling).
 */
typedef MkAdjT<Real> MkAdj;

extern "C" __global__ void __launch_bounds__(MK_BT)
mk_adj_vjp(const MkArgs a, const MkAdj j, const int s)
{
    const long long c = (long long)blockIdx.x * MK_BT + threadIdx.x;
    if (c >= a.n) return;
    const MkStep p = a.steps[s];
    Real y[MK_NS], lam[MK_NS], mu[MK_NS], fv[MK_NFA], fb[MK_NFA], ab;
#pragma unroll
    for (int i = 0; i < MK_NS; ++i) {
        y[i] = j.yn[(long long)i * a.npad + c];
        lam[i] = j.lam[(long long)i * a.npad + c];
    }
#pragma unroll
    for (int k = 0; k < MK_NF; ++k) fv[k] = a.fields[(long long)k * a.npad + c];
    const long long ix = c % a.nx, iy = c / a.nx;
    const Real* __restrict__ V = j.yn + (long long)MK_IVM * a.npad;
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
    mk_cell_vjp(y, p.t, p.dt, pace, idiff, fv, lam, mu, &ab, fb);
#pragma unroll
    for (int i = 0; i < MK_NS; ++i) j.mu[(long long)i * a.npad + c] = mu[i];
    j.abar[c] = ab;
#pragma unroll
    for (int k = 0; k < MK_NF; ++k)
        j.gfield[(long long)k * a.npad + c] += (double)fb[k];
}

extern "C" __global__ void __launch_bounds__(MK_BT)
mk_adj_couple(const MkArgs a, const MkAdj j, const int s)
{
    const long long c = (long long)blockIdx.x * MK_BT + threadIdx.x;
    if (c >= a.n) return;
    (void)s;
#pragma unroll
    for (int i = 0; i < MK_NS; ++i)
        j.lam_out[(long long)i * a.npad + c] =
            j.mu ? j.mu[(long long)i * a.npad + c] : (Real)0;
    const Real* __restrict__ V = j.yn + (long long)MK_IVM * a.npad;
    const Real v = V[c];
    Real lv = (Real)0;
    if (j.mu) {
        const Real ac = j.abar[c];
#if MK_COUPLING == 1
        const long long ix = c % a.nx, iy = c / a.nx;
        Real ax = (Real)0, xs = (Real)0, ay = (Real)0, ys = (Real)0;
        if (a.nx > 1) {
            if (ix > 0) { ax += ac - j.abar[c - 1]; xs += v - V[c - 1]; }
            if (ix < a.nx - 1) { ax += ac - j.abar[c + 1]; xs += v - V[c + 1]; }
        }
        if (a.ny > 1) {
            if (iy > 0) { ay += ac - j.abar[c - a.nx]; ys += v - V[c - a.nx]; }
            if (iy < a.ny - 1) {
                ay += ac - j.abar[c + a.nx];
                ys += v - V[c + a.nx];
            }
        }
        lv = a.gx * ax + a.gy * ay;
        j.gx_acc[c] += (double)ac * (double)xs;
        j.gy_acc[c] += (double)ac * (double)ys;
#elif MK_COUPLING == 2
        const long long ix = c % a.nx, iy = c / a.nx;
        if (a.nx > 1) {
            if (ix > 0) lv += a.gx_field[c - iy - 1] * (ac - j.abar[c - 1]);
            if (ix < a.nx - 1) lv += a.gx_field[c - iy] * (ac - j.abar[c + 1]);
        }
        if (a.ny > 1) {
            if (iy > 0) lv += a.gy_field[c - a.nx] * (ac - j.abar[c - a.nx]);
            if (iy < a.ny - 1) lv += a.gy_field[c] * (ac - j.abar[c + a.nx]);
        }
#elif MK_COUPLING == 3
        double gs = 0;
        for (int e = a.csr_ptr[c]; e < a.csr_ptr[c + 1]; ++e) {
            int nb = a.csr_idx[e];
            if (nb < 0) nb = ~nb;
            const Real g = a.csr_g[e];
            lv += g * (ac - j.abar[nb]);
            gs += (double)g * ((double)ac - (double)j.abar[nb])
                  * ((double)v - (double)V[nb]);
        }
        /* each connection is seen from both ends: half each */
        j.gx_acc[c] += 0.5 * gs;
#endif
    }
    if (j.target_row >= 0 && j.probe_of[c] >= 0) {
        const int q = j.probe_of[c];
        const double d = (double)v - j.target[(long long)j.target_row
                                              * j.n_probe + q];
        lv += (Real)(2.0 * d);
        j.loss_acc[q] += d * d;
    }
    j.lam_out[(long long)MK_IVM * a.npad + c] += lv;
}

/* V at the probe cells after step s (trace row s). */
extern "C" __global__ void __launch_bounds__(64)
mk_adj_probe(const MkArgs a, const MkAdj j, const int s)
{
    const int q = blockIdx.x * 64 + threadIdx.x;
    if (q >= j.n_probe) return;
    j.trace[(long long)s * j.n_probe + q] =
        (double)j.yn[(long long)MK_IVM * a.npad + j.probes[q]];
}
