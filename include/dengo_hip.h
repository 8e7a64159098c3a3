/*
 * C ABI of the MI355X backend for Dengo's backward-Euler chemistry solver.
 *
 * Drop-in entry points (declared in primordial_solver.h, same names, argument
 * order and meaning as the emitted solver):
 *   primordial_setup_data      <- rates_and_rate_tables.C.template:17-54
 *   primordial_read_rate_tables / primordial_read_cooling_tables
 *                              <- rates_and_rate_tables.C.template:278-304
 *   dengo_evolve_primordial    <- rates_and_rate_tables.C.template:149-274
 *                                 (double &dt becomes double *dt)
 *   calculate_rhs_primordial   <- rates_and_rate_tables.C.template:460-522
 *   calculate_jacobian_primordial <- rates_and_rate_tables.C.template:527-599
 *   primordial_calculate_temperature <- rates_and_rate_tables.C.template:309-350
 *
 * Semantics of the drop-in: every cell is integrated as its own strip of one
 * cell (dengo_evolve with nstrip = 1): it keeps its own dt, its own Newton
 * convergence test and its own electron total.  The return value is the
 * smallest elapsed time over cells (< dtf means some cell did not finish, the
 * reference's only failure signal) and *dt is that cell's final substep.
 * Tables are read from <name>_tables.bin in the working directory (a 64-byte
 * header "DENGOTB1", int32 ntables, int32 nT, double T0, T1, z0, z1, then
 * FP64 [ntables][nT]); the reference reads <name>_tables.h5.
 *
 * Backend extensions are declared below; every function returns 0 on success
 * and a negative code on failure (message from dengo_last_error()).
 */
#ifndef DENGO_HIP_H
#define DENGO_HIP_H

#include "primordial_solver.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic of the Newton linear solve and schedule of cells on the GPU. */
enum dengo_precision {
    DENGO_FP64 = 0,       /* reference-faithful FP64: same operation order and
                             rounding as a plain C build of the algorithm */
    DENGO_MIXED = 1,      /* FP32 LU of the row-equilibrated species block,
                             FP64 residual, state and energy row; the Newton
                             loop re-evaluates the FP64 residual every sweep */
    DENGO_FP64_FAST = 2   /* FP64 block solve exploiting the network's
                             structure (energy column decoupled, structural
                             zeros skipped, reciprocal pivots, FMA) */
};
enum dengo_schedule {
    DENGO_PER_CELL = 0,   /* one cell per lane, grid covers all cells */
    DENGO_PERSISTENT = 1  /* resident lanes pull cells from a device queue */
};

typedef struct dengo_options {
    int precision;        /* enum dengo_precision */
    int schedule;         /* enum dengo_schedule */
    int block;            /* threads per block, multiple of 64 (0: default) */
    int grid;             /* blocks for the persistent schedule (0: fill GPU) */
    int max_substeps;     /* accepted-substep cap, reference: 50000 */
    int input_is_number;  /* 1: species given as number densities */
    int reserved[6];
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

void dengo_default_options(dengo_options *opt);
const char *dengo_last_error(void);

/* Allocate an empty data struct (no file I/O). */
primordial_data *primordial_alloc_data(void);
/* Set the tables from memory: tables[t * nT + i] in struct order, T grid
 * log-spaced on [T0, T1].  Marks the device copy stale. */
int primordial_set_tables(primordial_data *data, const double *tables, int ntables,
                          int nT, double T0, double T1);
/* As primordial_setup_data, reading the table blob from an explicit path. */
primordial_data *primordial_setup_data_from(const char *path, int *NumberOfFields,
                                            char ***FieldNames);

/* Integrate `dims` cells held in device memory (AoS [dims][10], updated in
 * place, as dengo_evolve does with host arrays), enqueued on `stream`
 * (a hipStream_t; NULL = default stream).  dt0 < 0 selects dtf/1e5.  When
 * ttot_min is non-NULL the call synchronises and returns the smallest elapsed
 * time there, and that cell's final dt in *dt_min. */
int dengo_primordial_evolve_device(primordial_data *data, const dengo_options *opt,
                                   double dtf, double dt0, double z, double *d_input,
                                   const double *d_rtol, double *d_atol, long long dims,
                                   const dengo_cell_outputs *outputs, double *ttot_min,
                                   double *dt_min, void *stream);

/* Host-array variant of dengo_evolve_primordial with options and host per-cell
 * outputs (any may be NULL). */
double dengo_primordial_evolve_host(primordial_data *data, const dengo_options *opt,
                                    double dtf, double *dt, double z, double *input,
                                    double *rtol, double *atol, long long dims,
                                    int *status, int *nsub, int *nrej, int *nsweep,
                                    double *dt_cell, double *T_cell);

/* Batched evaluation of the RHS f[dims][10], Jacobian J[dims][10][10]
 * (column-major per cell, may be NULL), temperature T[dims] (may be NULL) and
 * interpolated tables rates[dims][45] (may be NULL) for number-density states
 * u[dims][10] on host memory.  Returns the number of cells whose RHS saw a
 * negative value, or a negative error code. */
long long dengo_primordial_eval(primordial_data *data, const double *u, long long dims,
                                double z, double *f, double *J, double *T, double *rates);

/* Kernel timing of the last evolve call on the library's own events (ms). */
double dengo_last_kernel_ms(void);

#ifdef __cplusplus
}
#endif
#endif
