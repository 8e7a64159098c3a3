/*
 * C ABI of the dense-Jacobian solver for the synthetic 128-species network
 * (isothermal cells, reactions a + b -> c + d with Arrhenius rates
 * k = A (T/300)^beta exp(-gamma/T), evaluated as A exp(beta ln(T/300) -
 * gamma/T)).
 *
 * The reference has no generated solver for this network; the entry points
 * keep the shape of the generated solver's (dengo/templates/
 * rates_and_rate_tables.h.template:19-91, rates_and_rate_tables.C.template:
 * 149-274) with the per-cell temperatures in place of the redshift, since
 * these cells carry no energy equation.  The integration semantics (substep
 * control, BE Newton sweeps, Gaussian elimination) are the reference's
 * (dengo/solvers/BE_chem_solve.C:79-382).  All states are number densities.
 */
#ifndef SYNTHETIC_SOLVER_H
#define SYNTHETIC_SOLVER_H

#include "dengo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SYNTHETIC_NSPECIES 128
#define SYNTHETIC_MAX_REACTIONS 2048

typedef struct synthetic_data synthetic_data;

/* Network from reaction rows reactions[4 r .. 4 r + 3] = (a, b, c, d) and
 * Arrhenius parameters; nspecies must be SYNTHETIC_NSPECIES.  Returns NULL on
 * failure (message in dengo_last_error()).  Counterpart of
 * <name>_setup_data (rates_and_rate_tables.C.template:17-54). */
synthetic_data *synthetic_setup_data(int nspecies, int nreactions, const int *reactions,
        const double *A, const double *beta, const double *gamma);
/* Releases the struct and its device buffers. */
void synthetic_free_data(synthetic_data *data);

/* Counterpart of dengo_evolve_<name> (rates_and_rate_tables.C.template:
 * 149-274) on host arrays input/rtol/atol[dims][128] (input and atol updated
 * in place), T[dims] in K; reference-faithful FP64.  Returns the smallest
 * elapsed time over the cells and sets *dt to that cell's last substep. */
double dengo_evolve_synthetic(double dtf, double *dt, const double *T, double *input,
        double *rtol, double *atol, long long dims, synthetic_data *data);

/* Device-memory form, enqueued on stream (a hipStream_t, NULL = default).
 * opt->precision: DENGO_FP64 (faithful) or DENGO_FP64_FAST (Gauss-Jordan,
 * FMA); opt->schedule: DENGO_PERSISTENT (workgroups pull cells from a device
 * queue) or DENGO_PER_CELL; opt->block must be 0 or 256.  dt0 < 0 selects
 * dtf/1e5.  With ttot_min non-NULL the call synchronises and returns the
 * smallest elapsed time and that cell's final dt. */
int dengo_synthetic_evolve_device(synthetic_data *data, const dengo_options *opt, double dtf,
        double dt0, const double *d_T, double *d_input, const double *d_rtol, double *d_atol,
        long long dims, const dengo_cell_outputs *outputs, double *ttot_min, double *dt_min,
        void *stream);

/* Host-array form with options and per-cell outputs (any may be NULL). */
double dengo_synthetic_evolve_host(synthetic_data *data, const dengo_options *opt, double dtf,
        double *dt, const double *T, double *input, double *rtol, double *atol, long long dims,
        int *status, int *nsub, int *nrej, int *nsweep, double *dt_cell);

/* RHS f[dims][128] and Jacobian J[dims][128][128] (column-major per cell,
 * J[c][col][row]; may be NULL) of number-density states u at temperatures
 * T; returns the number of cells with a negative density (their f and J are
 * left untouched), or a negative error code. */
long long dengo_synthetic_eval(synthetic_data *data, const double *T, const double *u,
        long long dims, double *f, double *J);

#ifdef __cplusplus
}
#endif
#endif
