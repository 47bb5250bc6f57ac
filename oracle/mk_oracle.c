/*
 * mk_oracle.c -- CPU ORACLE (test infrastructure, not part of the product).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this code, and only as the checker or the CPU baseline.
 *
 * A plain-C, double-precision restatement of what myokit's SimulationOpenCL
 * computes, written from the reference sources:
 *   - time-step and log rule of sim_step       myokit/_sim/openclsim.c:1051-1162
 *   - order within a step (diffusion at t, NaN check, cell step, log,
 *     time and pacing advance)                  openclsim.c:1066-1155
 *   - homogeneous stencil diff_step             myokit/_sim/openclsim.cl:384-435
 *   - conductance-field stencil diff_hetero     openclsim.cl:452-487
 *   - connections diff_arb_reset/diff_arb_step  openclsim.cl:537-573
 *     (atomic adds taken in connection-list order)
 *   - per-cell pacing calculate_pacing          openclsim.cl:249-280
 *   - FE / RL update                            openclsim.cl:358-364
 *   - event-based pacing ESys                   myokit/_sim/pacing.h:168-548
 * The model's equations come from oracle_cell(), generated per model by
 * oracle/cgen.py. Transcendentals use myokit_amd/csrc/mk_math.h, so a
 * double-precision device run can be compared bit for bit.
 *
 * Cells are independent within a step, so the cell loop is parallelised
 * with OpenMP; results do not depend on the thread count.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Provided by the generated model file */
extern const int oracle_n_state;
extern const int oracle_n_field;
extern const int oracle_n_quant;
void oracle_cell(double* y, double t, double dt, double pace, double idiff,
                 const double* fv, double* q, int log);

/* ---------------------------------------------------------------- ESys */
typedef struct {
    double level, duration, start, period, multiplier;
    int next;
} OEvent;

typedef struct {
    OEvent* ev;
    int n, head, fire;
    double time, tnext, tdown, level;
} OESys;

static int o_eq(double a, double b)
{
    double s;
    if (a == b) return 1;
    s = fabs(a) > fabs(b) ? fabs(a) : fabs(b);
    return fabs(a - b) / s < DBL_EPSILON;
}

static int o_geq(double a, double b) { return a >= b || o_eq(a, b); }

/* Insert event k into the queue; returns 0 or -50 (simultaneous). */
static int o_schedule(OESys* s, int k)
{
    int e, flag = 0;
    OEvent* add = &s->ev[k];
    add->next = -1;
    if (s->head < 0) { s->head = k; return 0; }
    if (add->start < s->ev[s->head].start) {
        add->next = s->head;
        s->head = k;
        return 0;
    }
    e = s->head;
    while (s->ev[e].next >= 0 && add->start >= s->ev[s->ev[e].next].start)
        e = s->ev[e].next;
    if (add->start == s->ev[e].start) flag = -50;
    add->next = s->ev[e].next;
    s->ev[e].next = k;
    return flag;
}

static int o_advance(OESys* s, double t)
{
    int r;
    if (t < s->time) return -40;
    s->time = t;
    while (o_geq(s->time, s->tnext)) {
        if (s->fire >= 0 && o_geq(s->tnext, s->tdown)) {
            s->fire = -1;
            s->level = 0;
        }
        if (s->head >= 0 && o_geq(s->tnext, s->ev[s->head].start)) {
            OEvent* f;
            s->fire = s->head;
            f = &s->ev[s->fire];
            s->head = f->next;
            s->tdown = f->start + f->duration;
            s->level = f->level;
            if (f->period > 0) {
                if (f->multiplier != 1) {
                    if (f->multiplier > 1) f->multiplier -= 1;
                    f->start += f->period;
                    r = o_schedule(s, s->fire);
                    if (r) return r;
                } else {
                    f->period = 0;
                }
            }
            if (s->head >= 0 && o_eq(s->ev[s->head].start, s->tdown))
                s->tdown = s->ev[s->head].start;
        }
        s->tnext = HUGE_VAL;
        if (s->fire >= 0 && s->tnext > s->tdown) s->tnext = s->tdown;
        if (s->head >= 0 && s->tnext > s->ev[s->head].start)
            s->tnext = s->ev[s->head].start;
    }
    return 0;
}

static int o_esys_init(OESys* s, double t0, const double* rows, long n)
{
    long k;
    int r;
    s->n = (int)n;
    s->ev = (OEvent*)calloc(n > 0 ? n : 1, sizeof(OEvent));
    s->head = s->fire = -1;
    s->time = s->tnext = s->tdown = t0;
    s->level = 0;
    for (k = 0; k < n; ++k) {
        s->ev[k].level = rows[5 * k + 0];
        s->ev[k].start = rows[5 * k + 1];
        s->ev[k].duration = rows[5 * k + 2];
        s->ev[k].period = rows[5 * k + 3];
        s->ev[k].multiplier = rows[5 * k + 4];
        s->ev[k].next = -1;
    }
    for (k = 0; k < n; ++k) {
        r = o_schedule(s, (int)k);
        if (r) return r;
    }
    return o_advance(s, t0);
}

/* ------------------------------------------------------------- the run */
typedef struct {
    long nx, ny;
    int coupling;               /* 0 none, 1 stencil, 2 hetero, 3 graph */
    int i_vm;
    double gx, gy;
    const double* gx_field;     /* (ny, nx-1) */
    const double* gy_field;     /* (ny-1, nx) */
    long n_conn;
    const long* conn_i;
    const long* conn_j;
    const double* conn_g;
    int pace_mode;              /* 0 all, 1 rectangle, 2 mask, 3 none */
    long px0, px1, py0, py1;
    const unsigned char* pace_mask;
    const double* fields;       /* cell-major, n_field per cell */
    double* state;              /* AoS in/out, n_state per cell */
    double tmin, tmax, dt;
    const double* events;
    long n_events;
    double log_interval;
    const int* colmap;          /* [n_quant][n] */
    int n_cols;
    double* log_values;         /* [max_rows][n_cols] */
    double* log_time;
    double* log_pace;
    long max_rows;
    int nan_check;
    int threads;
    /* outputs */
    long rows;
    double t_end;
    int halted;
    long steps;
} OracleArgs;

static int o_paced(const OracleArgs* a, long c, long ix, long iy)
{
    if (a->pace_mode == 0) return 1;
    if (a->pace_mode == 1)
        return ix >= a->px0 && ix < a->px1 && iy >= a->py0 && iy < a->py1;
    if (a->pace_mode == 2) return a->pace_mask[c] != 0;
    return 0;
}

int oracle_run(OracleArgs* a)
{
    const long n = a->nx * a->ny;
    const int ns = oracle_n_state, nf = oracle_n_field, nq = oracle_n_quant;
    double* idiff = (double*)calloc(n > 0 ? n : 1, sizeof(double));
    OESys es;
    double t = a->tmin, tnext_log = a->tmin, tnext_pace, pace;
    unsigned long istep = 1, inext_log = 0;
    long rows = 0, steps = 0;
    int r, halt = 0;

    r = o_esys_init(&es, a->tmin, a->events, a->n_events);
    if (r) { free(idiff); free(es.ev); return r; }
    tnext_pace = es.tnext;
    pace = es.level;

    for (;;) {
        const int logging = t >= tnext_log;
        int inter = 0;
        double dt = a->tmin + (double)istep * a->dt - t, d;
        long c;
        d = a->tmax - t; if (d > 0 && d < dt) { dt = d; inter = 1; }
        d = tnext_pace - t; if (d > 0 && d < dt) { dt = d; inter = 1; }
        d = tnext_log - t; if (d > 0 && d < dt) { dt = d; inter = 1; }
        if (!inter) istep++;

        /* Diffusion current at time t */
        if (a->coupling == 1) {
            #pragma omp parallel for num_threads(a->threads) schedule(static)
            for (c = 0; c < n; ++c) {
                const long ix = c % a->nx, iy = c / a->nx;
                const double* s = a->state;
                const long of1 = c * ns + a->i_vm;
                double i;
                if (a->nx > 1) {
                    if (ix == 0) i = a->gx * (s[of1] - s[of1 + ns]);
                    else if (ix == a->nx - 1) i = a->gx * (s[of1] - s[of1 - ns]);
                    else i = a->gx * (2 * s[of1] - s[of1 - ns] - s[of1 + ns]);
                } else {
                    i = 0;
                }
                if (a->ny > 1) {
                    const long up = of1 + ns * a->nx, dn = of1 - ns * a->nx;
                    if (iy == 0) i += a->gy * (s[of1] - s[up]);
                    else if (iy == a->ny - 1) i += a->gy * (s[of1] - s[dn]);
                    else i += a->gy * (2 * s[of1] - s[dn] - s[up]);
                }
                idiff[c] = i;
            }
        } else if (a->coupling == 2) {
            #pragma omp parallel for num_threads(a->threads) schedule(static)
            for (c = 0; c < n; ++c) {
                const long ix = c % a->nx, iy = c / a->nx;
                const double* s = a->state;
                const long off = c * ns + a->i_vm;
                const double v = s[off];
                double i = 0.0;
                if (a->nx > 1) {
                    if (ix > 0) i += a->gx_field[c - iy - 1] * (v - s[off - ns]);
                    if (ix < a->nx - 1) i += a->gx_field[c - iy] * (v - s[off + ns]);
                }
                if (a->ny > 1) {
                    if (iy > 0) i += a->gy_field[c - a->nx] * (v - s[off - ns * a->nx]);
                    if (iy < a->ny - 1) i += a->gy_field[c] * (v - s[off + ns * a->nx]);
                }
                idiff[c] = i;
            }
        } else if (a->coupling == 3) {
            long k;
            for (c = 0; c < n; ++c) idiff[c] = 0;
            for (k = 0; k < a->n_conn; ++k) {
                const long i1 = a->conn_i[k], i2 = a->conn_j[k];
                const double i12 = a->conn_g[k] *
                    (a->state[i1 * ns + a->i_vm] - a->state[i2 * ns + a->i_vm]);
                idiff[i1] = idiff[i1] + i12;
                idiff[i2] = idiff[i2] + (-i12);
            }
        }

        /* NaN test on the state at t (openclsim.c:1079-1090) */
        if (logging && a->nan_check && isnan(a->state[0])) halt = 1;

        /* Cell step: intermediaries at t, state to t + dt */
        {
            const long row = logging ? rows : -1;
            const int do_log = logging && rows < a->max_rows;
            #pragma omp parallel for num_threads(a->threads) schedule(static)
            for (c = 0; c < n; ++c) {
                const long ix = c % a->nx, iy = c / a->nx;
                double q[64];
                const double p = o_paced(a, c, ix, iy) ? pace : 0.0;
                oracle_cell(a->state + c * ns, t, dt, p,
                            a->coupling ? idiff[c] : 0.0,
                            nf ? a->fields + c * nf : NULL, q, do_log);
                if (do_log) {
                    int k;
                    for (k = 0; k < nq; ++k) {
                        const int col = a->colmap[(long)k * n + c];
                        if (col >= 0) a->log_values[row * a->n_cols + col] = q[k];
                    }
                }
            }
        }
        if (logging) {
            if (rows < a->max_rows) {
                a->log_time[rows] = t;
                a->log_pace[rows] = pace;
            }
            rows++;
            inext_log++;
            tnext_log = a->tmin + (double)inext_log * a->log_interval;
        }
        t += dt;
        steps++;
        r = o_advance(&es, t);
        if (r) { free(idiff); free(es.ev); return r; }
        tnext_pace = es.tnext;
        pace = es.level;
        if (t >= a->tmax || halt) break;
    }
    a->rows = rows;
    a->t_end = t;
    a->halted = halt;
    a->steps = steps;
    free(idiff);
    free(es.ev);
    return 0;
}

/* ------------------------------------------------- fiber-tissue run
 * FiberTissueSimulation's sim_step (myokit/_sim/fiber_tissue.c:1003-1160):
 * per step, the homogeneous stencil of each grid (diff_step) at t, the
 * junction current of diff_step_fiber_tissue (openclsim.cl:601-628) added
 * to the fiber cell and subtracted from the tissue cell, then the cell
 * step of every fiber cell (paced where x < nx_paced: the reference's
 * explicit cell list, fiber_tissue.py:195-198) and every tissue cell
 * (unpaced: `Real pace = 0`, openclsim.cl:324-329). The two cell models
 * come in as function pointers; states are AoS per part. Logs V of every
 * cell of both parts at the log points (fiber cells first). */
typedef void (*OracleCellFn)(double*, double, double, double, double,
                             const double*, double*, int);

typedef struct {
    OracleCellFn cell_f, cell_t;
    int nsf, nst, ivf, ivt;
    long nfx, nfy, ntx, nty, ctx, cty, nx_paced;
    double gfx, gfy, gtx, gty, gft;
    double* state_f;
    double* state_t;
    double tmin, tmax, dt;
    const double* events;
    long n_events;
    double log_interval;
    double* log_v;          /* [max_rows][nfx*nfy + ntx*nty] */
    double* log_time;
    double* log_pace;
    long max_rows;
    /* outputs */
    long rows;
    double t_end;
    long steps;
} OracleFtArgs;

/* diff_step (openclsim.cl:384-435) on an AoS grid */
static void o_stencil(const double* s, int ns, int iv, long nx, long ny,
                      double gx, double gy, double* idiff)
{
    long c;
    for (c = 0; c < nx * ny; ++c) {
        const long ix = c % nx, iy = c / nx;
        const long of1 = c * ns + iv;
        double i;
        if (nx > 1) {
            if (ix == 0) i = gx * (s[of1] - s[of1 + ns]);
            else if (ix == nx - 1) i = gx * (s[of1] - s[of1 - ns]);
            else i = gx * (2 * s[of1] - s[of1 - ns] - s[of1 + ns]);
        } else {
            i = 0;
        }
        if (ny > 1) {
            const long up = of1 + ns * nx, dn = of1 - ns * nx;
            if (iy == 0) i += gy * (s[of1] - s[up]);
            else if (iy == ny - 1) i += gy * (s[of1] - s[dn]);
            else i += gy * (2 * s[of1] - s[dn] - s[up]);
        }
        idiff[c] = i;
    }
}

int oracle_ft_run(OracleFtArgs* a)
{
    const long nf = a->nfx * a->nfy, nt = a->ntx * a->nty;
    double* idf = (double*)calloc(nf, sizeof(double));
    double* idt = (double*)calloc(nt, sizeof(double));
    OESys es;
    double t = a->tmin, tnext_log = a->tmin, tnext_pace, pace;
    unsigned long istep = 1, inext_log = 0;
    long rows = 0, steps = 0, c, k;
    int r;

    r = o_esys_init(&es, a->tmin, a->events, a->n_events);
    if (r) { free(idf); free(idt); free(es.ev); return r; }
    tnext_pace = es.tnext;
    pace = es.level;
    for (;;) {
        const int logging = t >= tnext_log;
        int inter = 0;
        double dt = a->tmin + (double)istep * a->dt - t, d;
        d = a->tmax - t; if (d > 0 && d < dt) { dt = d; inter = 1; }
        d = tnext_pace - t; if (d > 0 && d < dt) { dt = d; inter = 1; }
        d = tnext_log - t; if (d > 0 && d < dt) { dt = d; inter = 1; }
        if (!inter) istep++;

        o_stencil(a->state_f, a->nsf, a->ivf, a->nfx, a->nfy, a->gfx, a->gfy,
                  idf);
        o_stencil(a->state_t, a->nst, a->ivt, a->ntx, a->nty, a->gtx, a->gty,
                  idt);
        for (k = 0; k < a->nfy; ++k) {
            const long iff = (a->nfx - 1) + k * a->nfx;
            const long ift = a->ctx + (a->cty + k) * a->ntx;
            const double i = a->gft * (a->state_f[iff * a->nsf + a->ivf]
                                       - a->state_t[ift * a->nst + a->ivt]);
            idf[iff] += i;
            idt[ift] -= i;
        }
        if (logging && rows < a->max_rows) {
            double* row = a->log_v + rows * (nf + nt);
            for (c = 0; c < nf; ++c) row[c] = a->state_f[c * a->nsf + a->ivf];
            for (c = 0; c < nt; ++c)
                row[nf + c] = a->state_t[c * a->nst + a->ivt];
            a->log_time[rows] = t;
            a->log_pace[rows] = pace;
        }
        for (c = 0; c < nf; ++c) {
            const long ix = c % a->nfx;
            a->cell_f(a->state_f + c * a->nsf, t, dt,
                      ix < a->nx_paced ? pace : 0.0, idf[c], NULL, NULL, 0);
        }
        for (c = 0; c < nt; ++c)
            a->cell_t(a->state_t + c * a->nst, t, dt, 0.0, idt[c], NULL, NULL,
                      0);
        if (logging) {
            rows++;
            inext_log++;
            tnext_log = a->tmin + (double)inext_log * a->log_interval;
        }
        t += dt;
        steps++;
        r = o_advance(&es, t);
        if (r) { free(idf); free(idt); free(es.ev); return r; }
        tnext_pace = es.tnext;
        pace = es.level;
        if (t >= a->tmax) break;
    }
    a->rows = rows;
    a->t_end = t;
    a->steps = steps;
    free(idf);
    free(idt);
    free(es.ev);
    return 0;
}
