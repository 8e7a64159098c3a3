/*
 * CPU FP64 oracle for the Dengo backward-Euler chemistry solver.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates the reference algorithm on
 * the CPU so that the GPU path can be checked against it; it is imported only
 * by tests/, by the smoke check and by bench.py's CPU-baseline leg.  Nothing
 * in the product (dengo_amd/) links, loads or calls it.
 *
 * Semantics: one dengo_evolve call per cell (nstrip = 1), following
 *   dengo/templates/rates_and_rate_tables.C.template:149-274 (substep driver),
 *   :309-350 (temperature), :362-455 (interpolation), :460-522 (RHS),
 *   :527-599 (Jacobian), :604-630 (electron consistency),
 *   dengo/solvers/BE_chem_solve.C:79-382 (scaled Newton + Gauss_Elim).
 */
#ifndef DENGO_ORACLE_H
#define DENGO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef int (*orc_rhs_f)(double *, double *, int, int, void *);
typedef int (*orc_jac_f)(double *, double *, int, int, void *);
/* Same signature as BE_chem_solve (dengo/solvers/BE_chem_solve.C:79-83). */
typedef int (*orc_be_f)(orc_rhs_f, orc_jac_f, double *u, double dt, double *rtol,
                        double *atol, int nstrip, int nchem, double *scaling,
                        void *sdata, double *u0, double *s, double *gu, double *Ju);

/* Primordial network: 10 unknowns, 45 tables (order of the emitted struct). */
#define ORC_PRIM_N 10
#define ORC_PRIM_NTAB 45

/* Table values of the primordial network on temperatures T[0..n), written
 * as out[t*n + i]; restates dengo/primordial_rates.py:47-293 and
 * dengo/primordial_cooling.py:34-316 with libm. */
void orc_primordial_tables(const double *T, int n, double *out);

/* The scaled Newton solver, restated (BE_chem_solve.C:79-321). */
int orc_BE_chem_solve(orc_rhs_f f, orc_jac_f J, double *u, double dt, double *rtol,
                      double *atol, int nstrip, int nchem, double *scaling,
                      void *sdata, double *u0, double *s, double *gu, double *Ju);
/* Gauss_Elim, restated (BE_chem_solve.C:329-382). */
int orc_gauss_elim(double *A, double *x, double *b, int n);

/* Generic mass-action network with linear cooling terms (ion-by-ion
 * networks): reaction rate law k * prod(n_slot), cooling term
 * sign * c * n_species * de^depow.  The energy unknown is last. */
#define ORC_MAXN 16
#define ORC_MAXTAB 64
#define ORC_MAXRXN 64
typedef struct {
    int table, nslot, slot[3], nprod, pc[4], ps[4];   /* up to 4 distinct products */
} orc_rxn;
typedef struct {
    int table, species, depow;
    double sign;
} orc_cool;
typedef struct orc_ma_net {
    int n, ntab, nrxn, ncool, de;      /* energy index is n - 1 */
    double W[ORC_MAXN], free_e[ORC_MAXN];
    int h2gamma[ORC_MAXN];             /* 1: gamma = 7/5, else 5/3 */
    orc_rxn rxn[ORC_MAXRXN];
    orc_cool cool[ORC_MAXTAB];
} orc_ma_net;

/* A network whose species totals and cooling sum come from functions: the
 * generated expressions of dengo_amd.codegen compiled for the host.  The
 * oracle supplies everything around them (temperature, interpolation,
 * redshift tables, the negative-species check, the division of the energy
 * row by the mass density, BE/Newton and the substep driver). */
typedef void (*orc_gen_fn)(const double *u, const double *rs, double T, double z, double *out);
typedef struct orc_gen_net {
    int n, ntab, de;                   /* energy index is n - 1 */
    double W[ORC_MAXN], free_e[ORC_MAXN];
    int h2gamma[ORC_MAXN];
    int zslot[ORC_MAXTAB];             /* >= 0: a redshift table */
    double z_bounds[2];
    int nz;                            /* points of the redshift grid */
    orc_gen_fn rhs;                    /* f[s] = species total, f[ge] = cooling sum */
    orc_gen_fn jac;                    /* J[col * n + row] of the same forms */
} orc_gen_net;

/* Per-call state handed to the RHS/Jacobian callbacks (nstrip must be 1). */
typedef struct orc_prim_data {
    double bounds[2];
    int nbins;
    double dbin, idbin, lb;
    const double *tables;   /* [45][nT] */
    int nT;
    double current_z;
    /* last temperature / interpolation state (from the last RHS call) */
    double T, logT, invT, dTs_ge, Tdef, dT;
    int bin_id;
    double rs[ORC_MAXTAB];
    long long rhs_calls;    /* one per Newton sweep */
    int flags;              /* ORC_FLAG_* */
    const orc_ma_net *net;  /* NULL: the primordial network */
    int n, ntab;
    const orc_gen_net *gen; /* non-NULL: a generated network */
} orc_prim_data;

/* Accepted-substep cap (default 50000, as the reference). */
void orc_set_max_substeps(int n);
/* Intermediate-solution records of the per-cell evolve (rates_and_rate_tables.C.template:
 * 212-241): record k of cell c at buf[(k * ncells + c) * (N + 3)]: the state before accepted
 * substep k + 1 (unknown order, number densities), the last sweep's T, the elapsed time and
 * the next dt; at most cap per cell.  NULL turns recording off. */
void orc_set_trajectory(double *buf, long long cap);

/* Take ln T with orc_plog instead of libm's log. */
#define ORC_FLAG_PORTABLE_LOG 1
double orc_plog(double x);

void orc_prim_setup(orc_prim_data *d, const double *tables, int nT, double T0, double T1);
int orc_prim_rhs(double *input, double *rhs, int nstrip, int nchem, void *sdata);
int orc_prim_jac(double *input, double *J, int nstrip, int nchem, void *sdata);
void orc_prim_temperature(orc_prim_data *d, const double *u);
void orc_prim_interpolate(orc_prim_data *d);

/* orc_prim_interpolate at given ln T values: out[i * ntab + t] for ntab
 * tables of nT bins on [T0, T1] (ntab <= ORC_MAXTAB). */
void orc_interp_logT(const double *tables, int ntab, int nT, double T0, double T1,
                     const double *logT, long long n, double *out);

/* Whole-step driver over `dims` independent cells.  `input`, `rtol`, `atol`
 * are [dims][10] as in dengo_evolve; input and atol are updated in place.
 * dt: in = initial substep (<0 selects dtf/1e5), out = final dt of the cell
 * with the smallest elapsed time.  Returns the smallest elapsed time.
 * Optional per-cell outputs (may be NULL): status (0 reached dtf, 1 gave up
 * at dt/dtf < 1e-15, 2 hit the 50000-substep cap), accepted and rejected
 * substeps, Newton sweeps, final dt, final temperature.
 * be: NULL selects orc_BE_chem_solve.  nthreads <= 0 uses OpenMP's default.
 * flags: ORC_FLAG_* (0 = libm log, as the reference). */
double orc_prim_evolve(double dtf, double *dt, double z, double *input,
                       double *rtol, double *atol, long long dims,
                       const double *tables, int nT, double T0, double T1,
                       int input_is_number, int *status, int *nsub, int *nrej,
                       int *nsweep, double *dt_out, double *T_out,
                       orc_be_f be, int nthreads, int flags);

/* Evaluate RHS (f[dims][10]), Jacobian (J[dims][10][10], column-major per
 * cell) and temperature for states u (number densities, ge) at redshift z.
 * Returns the number of cells whose RHS reported a negative value. */
int orc_prim_eval(const double *u, long long dims, const double *tables, int nT,
                  double T0, double T1, double z, double *f, double *J,
                  double *T, double *rates, int flags);

double orc_pexp(double x);

/* Large synthetic network: rxn[4 r .. 4 r + 3] = (a, b, c, d) for
 * a + b -> c + d, Arrhenius k = A (T/300)^beta exp(-gamma/T). */
typedef struct orc_big_net {
    int n, nrxn;
    const int *rxn;
    const double *A, *beta, *gamma;
} orc_big_net;
void orc_big_rates(const orc_big_net *net, double T, double *k);
/* One isothermal BE integration per cell (number densities, no energy
 * equation); per-cell outputs may be NULL. */
double orc_big_evolve(const orc_big_net *net, double dtf, double *dt, const double *Tcell,
                      double *input, double *rtol, double *atol, long long dims, int *status,
                      int *nsub, int *nrej, int *nsweep, double *dt_out, int nthreads);

long long orc_big_eval(const orc_big_net *net, const double *Tcell, const double *u,
                       long long dims, double *f, double *J);

/* temperature_from_mass_density of nstrip cells of mass densities (net NULL:
 * the primordial network): T clamped below at 1 K, no upper clamp. */
void orc_temperature_from_mass_density(const orc_ma_net *net, const double *input, int nstrip,
                                       int nchem, double *T);

/* Value at redshift z of a table r on nz points log-spaced in 1 + z over
 * [z0, z1] (rates_and_rate_tables.C.template:384-401 and the pi/ph branches
 * of :402-449): linear in ln(1 + z), 0 outside z0 <= z < z1. */
double orc_z_rate(const double *r, int nz, double z0, double z1, double z);

/* Generated-network counterparts of orc_prim_evolve / orc_prim_eval; tables
 * [ntab][nT], a redshift table's first nz values in its row. */
double orc_gen_evolve(const orc_gen_net *net, double dtf, double *dt, double z, double *input,
                      double *rtol, double *atol, long long dims, const double *tables,
                      int nT, double T0, double T1, int input_is_number, int *status,
                      int *nsub, int *nrej, int *nsweep, double *dt_out, double *T_out,
                      int nthreads, int flags);
int orc_gen_eval(const orc_gen_net *net, const double *u, long long dims, double z,
                 const double *tables, int nT, double T0, double T1, double *f, double *J,
                 double *T, double *rates, int flags);

/* Mass-action network counterparts of orc_prim_evolve / orc_prim_eval. */
/* One BE_chem_solve (be; NULL selects orc_BE_chem_solve) of a strip of
 * nstrip cells of the primordial network (net NULL) or a mass-action
 * network, as the Cython harness calls it: u, atol [nstrip][n] in place,
 * scaling [nstrip][n]; T_out[nstrip] receives each cell's temperature of the
 * last RHS evaluation.  Returns the solver's 0 / 1. */
int orc_strip_be(const orc_ma_net *net, const double *tables, int nT, double T0, double T1,
                 double z, int flags, double *u, double dt, double *rtol, double *atol,
                 int nstrip, double *scaling, double *T_out, orc_be_f be);
/* ensure_electron_consistency on a strip (the total runs across cells). */
void orc_strip_electrons(const orc_ma_net *net, double *input, int nstrip);

double orc_ma_evolve(const orc_ma_net *net, double dtf, double *dt, double z, double *input,
                     double *rtol, double *atol, long long dims, const double *tables,
                     int nT, double T0, double T1, int input_is_number, int *status,
                     int *nsub, int *nrej, int *nsweep, double *dt_out, double *T_out,
                     orc_be_f be, int nthreads, int flags);
int orc_ma_eval(const orc_ma_net *net, const double *u, long long dims, const double *tables,
                int nT, double T0, double T1, double *f, double *J, double *T,
                double *rates, int flags);

#ifdef __cplusplus
}
#endif
#endif
