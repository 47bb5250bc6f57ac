/*
 * mk_args.h -- argument block shared by the host library and the generated
 * kernels (included by both; identical layout on host and device).
 *
 * Device layout: structure-of-arrays. State variable i of cell c lives at
 * state[i * npad + c], field k at fields[k * npad + c]. Cells are numbered
 * c = x + y * nx, as in the reference (openclsim.cl:316).
 */
#ifndef MK_ARGS_H
#define MK_ARGS_H

template <typename R>
struct MkStepT {
    R t;        /* engine time, cast to the simulation precision */
    R dt;       /* step size, cast to the simulation precision */
    R pace;     /* pacing level */
    int row;    /* device log row written at this step, or -1 */
};

template <typename R>
struct MkArgsT {
    R* state_a;                 /* SoA state, buffer A */
    R* state_b;                 /* SoA state, buffer B (grid driver) */
    const R* fields;            /* SoA per-cell fields */
    const MkStepT<R>* steps;    /* step table of the current batch */
    R* logbuf;                  /* [rows][ncols] per-cell log values */
    const int* colmap;          /* [nquant][npad] -> column or -1 */
    const unsigned char* pace_mask;  /* [npad] 1 = paced (pace_mode 2) */
    const int* csr_ptr;         /* graph: [n + 1] row pointers */
    const int* csr_idx;         /* graph: neighbour j, or ~j when the cell is
                                   the second cell of the connection */
    const R* csr_g;             /* graph: conductance per entry */
    const R* gx_field;          /* hetero: (ny, nx - 1) */
    const R* gy_field;          /* hetero: (ny - 1, nx) */
    const R* idiff_extra;       /* grid driver: current added to idiff where
                                   nonzero (fiber-tissue junction), or NULL */
    const long long* dom_off;   /* block driver: domain offsets into dom_cells */
    const int* dom_cells;       /* block driver: cell ids, grouped by domain
                                   (NULL = identity order) */
    int* halt;                  /* halt[0]: log row where NaN was found */
    long long n;                /* number of cells */
    long long npad;             /* padded cell count (SoA stride) */
    long long nx, ny;           /* grid size */
    long long px0, px1, py0, py1;   /* paced rectangle [px0,px1) x [py0,py1) */
    long long nan_cell;         /* cell whose state 0 the NaN test reads: the
                                   caller's cell 0 (openclsim.c:1087-1089) in
                                   the device's cell order */
    int pace_mode;              /* 0 all cells, 1 rectangle, 2 mask, 3 none */
    int ncols;                  /* per-cell log columns */
    int nan_check;              /* test state 0 of nan_cell at log steps */
    int ndom;                   /* number of domains (block driver) */
    int gcopies;                /* graph of identical copies interleaved:
                                   CSR rows are copy positions (device cell
                                   c is row c / gcopies), neighbours are
                                   positions (q -> q * gcopies + c %
                                   gcopies); 1: a row per cell */
    R gx, gy;                   /* homogeneous conductances */
};

/* Discrete adjoint of one step (see mk_adjoint.hip). */
template <typename R>
struct MkAdjT {
    const R* yn;            /* y_s, state before step s (SoA); NULL in the
                               closing pass */
    const R* yn1;           /* y_{s+1} = F(y_s) (SoA) */
    const R* mu_in;         /* (dF_{s+1}/dy)^T lam_{s+2} per cell; NULL = 0 */
    const R* ab_in;         /* adjoint of idiff at step s+1; NULL = 0 */
    R* mu_out;              /* (dF_s/dy)^T lam_{s+1} per cell (SoA) */
    R* ab_out;              /* adjoint of idiff at step s */
    double* gfield;         /* [n_field][npad] field gradients (accumulated) */
    double* gx_acc;         /* [workgroups] d/dgx partial sums (accumulated) */
    double* gy_acc;         /* [workgroups] d/dgy partial sums (accumulated) */
    double* loss_acc;       /* [n_probe] squared errors (accumulated) */
    const int* probe_of;    /* [npad] probe index of each cell, or -1 */
    const double* target;   /* [rows][n_probe] target potentials */
    const long long* probes;/* [n_probe] probe cells */
    double* trace;          /* [rows][n_probe] V at the probes, or NULL */
    int n_probe;
    int target_row;         /* target row compared with V(y_{s+1}), or -1 */
};


/* Persistent sheet driver: neighbour exchange between workgroup tiles.
 * An exchanged value is one 64-bit word per 32 bits of the value, each word
 * holding the step it belongs to (high half) and 32 bits of the value (low
 * half). */
template <typename R>
struct MkSheetT {
    unsigned long long* edges;  /* [2][tiles][4][64][sizeof(R) / 4] tile
                                   edge values, tagged (by step parity) */
    int* err;         /* set to 1 if a neighbour wait timed out */
    int tiles_x;      /* tiles along x (64 cells each) */
    int tiles_y;      /* tiles along y (MK_STH cells each) */
    int g0;           /* steps completed before this launch */
    int spin_limit;   /* polls before a wait is abandoned */
    R* traj;          /* if set: the state after local step i is stored at
                         traj + i * n_state * npad (SoA), for the adjoint */
};

/* Persistent chain driver: cells of disjoint chains (cables) laid end to
 * end, tile_len (at most 64 * MK_CCPT) consecutive cells per one-wave
 * workgroup. */
template <typename R>
struct MkChainT {
    const R* gl;              /* [npad] conductance to cell c - 1 */
    const R* gr;              /* [npad] conductance to cell c + 1 */
    const unsigned char* cf;  /* [npad] bit 0: connected to c - 1, bit 1:
                                 to c + 1, bit 2: the c - 1 connection comes
                                 first in the cell's connection-list order */
    unsigned long long* edges;  /* [2][tiles][2][sizeof(R) / 4] first and
                                   last V of each tile, tagged as the
                                   sheet's (by step parity) */
    int* err;                 /* set to 1 if a neighbour wait timed out */
    int tiles;                /* number of tiles */
    int tile_len;             /* cells per tile (the last may have fewer) */
    int g0;                   /* steps completed before this launch */
    int spin_limit;           /* polls before a wait is abandoned */
};

#endif
