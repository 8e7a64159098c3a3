/*
 * Types and network-independent entry points of the MI355X backend for
 * Dengo's backward-Euler chemistry solver.  Every function returns 0 on
 * success and a negative code on failure (message from dengo_last_error()).
 */
#ifndef DENGO_TYPES_H
#define DENGO_TYPES_H

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic of the Newton sweep. */
enum dengo_precision {
    DENGO_FP64 = 0,       /* reference-faithful FP64: same operation order and
                             rounding as a plain C build of the algorithm */
    DENGO_MIXED = 1,      /* FP32 LU of the equilibrated species block with
                             iterative refinement against the FP64 block; FP64
                             residual, state and energy row */
    DENGO_FP64_FAST = 2   /* FP64 block solve exploiting the network's
                             structure (energy column decoupled, structural
                             zeros skipped, reciprocal pivots, FMA) */
};
/* Schedule of cells on the GPU. */
enum dengo_schedule {
    DENGO_PER_CELL = 0,   /* one cell per lane, grid covers all cells */
    DENGO_PERSISTENT = 1  /* resident lanes pull cells from a device queue */
};

/* A network's hooks for the patch server (opaque; each solver library's
   dengo_<n>_server_hooks returns its own, for dengo_server_add_network). */
typedef struct dengo_net_hooks dengo_net_hooks;

typedef struct dengo_options {
    int precision;        /* enum dengo_precision */
    int schedule;         /* enum dengo_schedule */
    int block;            /* threads per block: 0 or 64 */
    int grid;             /* blocks for the persistent schedule (0: fill GPU) */
    int max_substeps;     /* accepted-substep cap, reference: 50000 */
    int input_is_number;  /* 1: species given as number densities */
    int kernel_form;      /* FP64-fast sweep: 0 the register-lean kernel where
                             the network has one that fits (two waves per
                             SIMD), else the register-resident one; 1 always
                             the register-resident one.  Both give the same
                             results bit for bit.  Other precisions ignore it. */
    int reserved[1];
    const long long *order;  /* NULL, or device memory holding the order in
                             which cells are handed out: cell order[0] first,
                             then order[1], ... (a permutation of 0..dims-1).
                             Handing out the costly cells first shortens the
                             tail of a persistent launch; a previous call's
                             per-cell sweep counts, sorted descending, are a
                             good order.  Results per cell do not depend on
                             it.  The evolve calls check that it is a
                             permutation (one pass, one synchronisation) and
                             fail otherwise; a captured graph checks it once,
                             at creation. */
    int reserved2[2];
} dengo_options;

/* Per-cell device outputs; any pointer may be NULL. */
typedef struct dengo_cell_outputs {
    int *status;          /* 0 reached dtf, 1 dt/dtf < 1e-15, 2 substep cap */
    int *nsub;            /* accepted substeps */
    int *nrej;            /* rejected substeps */
    int *nsweep;          /* Newton sweeps (RHS evaluations) */
    double *dt;           /* final substep */
    double *ttot;         /* elapsed time */
    double *T;            /* temperature of the last RHS evaluation */
} dengo_cell_outputs;

/* Backend extension (no counterpart in the reference): default options. */
void dengo_default_options(dengo_options *opt);
/* Backend extension: message of the last failure on this thread. */
const char *dengo_last_error(void);
/* Backend extension: kernel time of the last synchronising evolve call, from
 * HIP events on the launch stream (ms). */
double dengo_last_kernel_ms(void);
/* Backend extension (no counterpart in the reference): the FP64 vector FMA
 * rate of the current device in TFLOP/s, from one timed launch of
 * independent v_fma_f64 chains (iters FMAs per chain, 16 chains per lane,
 * eight 256-thread workgroups per CU).  The lane-per-cell sweeps' compute
 * ceiling, measured beside the published peak.  Allocates and frees its own
 * buffer; synchronises. */
int dengo_measure_fp64_fma_rate(int iters, double *tflops);

#ifdef __cplusplus
}
#endif
#endif
