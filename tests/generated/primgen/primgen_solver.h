/* Drop-in header for the primgen network (generated by dengo_amd.codegen).
   Struct fields and order follow dengo/templates/rates_and_rate_tables.h.template:30-83;
   entry points follow :85-108 with C linkage. */
#ifndef PRIMGEN_SOLVER_H
#define PRIMGEN_SOLVER_H

#ifndef MAX_NCELLS
#define MAX_NCELLS 1024
#endif
#define PRIMGEN_NSPECIES 10
#define PRIMGEN_NTABLES 45
#define PRIMGEN_NT 1024
/* Temperature grid of the tables and <name>_main's dtf
   (rates_and_rate_tables.C.template:25-26, 110). */
#define PRIMGEN_T_LO 1.0
#define PRIMGEN_T_HI 1000000000.0
#define PRIMGEN_STOP_TIME 1000000000000.0
/* ChemicalNetwork(write_intermediate=...): dengo_evolve writes
   <name>_intermediate_%%06d.h5 after every accepted substep
   (rates_and_rate_tables.C.template:212-241). */
#define PRIMGEN_WRITE_INTERMEDIATE 0
/* Redshift grid of the pi/ph tables (rates_and_rate_tables.C.template:31-36). */
#define PRIMGEN_NZ 1
#define PRIMGEN_Z_LO 0.0
#define PRIMGEN_Z_HI 0.0

typedef struct primgen_data_ {
    double dbin;
    double idbin;
    double bounds[2];
    int nbins;
    double d_zbin;
    double id_zbin;
    double z_bounds[2];
    int n_zbins;
    double current_z;
    double zdef;
    double dz;
    double Ts[MAX_NCELLS];
    double Tdef[MAX_NCELLS];
    double dT[MAX_NCELLS];
    double logTs[MAX_NCELLS];
    double invTs[MAX_NCELLS];
    double dTs_ge[MAX_NCELLS];
    double r_k01[1024];
    double rs_k01[MAX_NCELLS];
    double drs_k01[MAX_NCELLS];
    double r_k02[1024];
    double rs_k02[MAX_NCELLS];
    double drs_k02[MAX_NCELLS];
    double r_k03[1024];
    double rs_k03[MAX_NCELLS];
    double drs_k03[MAX_NCELLS];
    double r_k04[1024];
    double rs_k04[MAX_NCELLS];
    double drs_k04[MAX_NCELLS];
    double r_k05[1024];
    double rs_k05[MAX_NCELLS];
    double drs_k05[MAX_NCELLS];
    double r_k06[1024];
    double rs_k06[MAX_NCELLS];
    double drs_k06[MAX_NCELLS];
    double r_k07[1024];
    double rs_k07[MAX_NCELLS];
    double drs_k07[MAX_NCELLS];
    double r_k08[1024];
    double rs_k08[MAX_NCELLS];
    double drs_k08[MAX_NCELLS];
    double r_k09[1024];
    double rs_k09[MAX_NCELLS];
    double drs_k09[MAX_NCELLS];
    double r_k10[1024];
    double rs_k10[MAX_NCELLS];
    double drs_k10[MAX_NCELLS];
    double r_k11[1024];
    double rs_k11[MAX_NCELLS];
    double drs_k11[MAX_NCELLS];
    double r_k12[1024];
    double rs_k12[MAX_NCELLS];
    double drs_k12[MAX_NCELLS];
    double r_k13[1024];
    double rs_k13[MAX_NCELLS];
    double drs_k13[MAX_NCELLS];
    double r_k14[1024];
    double rs_k14[MAX_NCELLS];
    double drs_k14[MAX_NCELLS];
    double r_k15[1024];
    double rs_k15[MAX_NCELLS];
    double drs_k15[MAX_NCELLS];
    double r_k16[1024];
    double rs_k16[MAX_NCELLS];
    double drs_k16[MAX_NCELLS];
    double r_k17[1024];
    double rs_k17[MAX_NCELLS];
    double drs_k17[MAX_NCELLS];
    double r_k18[1024];
    double rs_k18[MAX_NCELLS];
    double drs_k18[MAX_NCELLS];
    double r_k19[1024];
    double rs_k19[MAX_NCELLS];
    double drs_k19[MAX_NCELLS];
    double r_k21[1024];
    double rs_k21[MAX_NCELLS];
    double drs_k21[MAX_NCELLS];
    double r_k22[1024];
    double rs_k22[MAX_NCELLS];
    double drs_k22[MAX_NCELLS];
    double r_k23[1024];
    double rs_k23[MAX_NCELLS];
    double drs_k23[MAX_NCELLS];
    double c_brem_brem[1024];
    double cs_brem_brem[MAX_NCELLS];
    double dcs_brem_brem[MAX_NCELLS];
    double c_ceHeI_ceHeI[1024];
    double cs_ceHeI_ceHeI[MAX_NCELLS];
    double dcs_ceHeI_ceHeI[MAX_NCELLS];
    double c_ceHeII_ceHeII[1024];
    double cs_ceHeII_ceHeII[MAX_NCELLS];
    double dcs_ceHeII_ceHeII[MAX_NCELLS];
    double c_ceHI_ceHI[1024];
    double cs_ceHI_ceHI[MAX_NCELLS];
    double dcs_ceHI_ceHI[MAX_NCELLS];
    double c_ciHeI_ciHeI[1024];
    double cs_ciHeI_ciHeI[MAX_NCELLS];
    double dcs_ciHeI_ciHeI[MAX_NCELLS];
    double c_ciHeII_ciHeII[1024];
    double cs_ciHeII_ciHeII[MAX_NCELLS];
    double dcs_ciHeII_ciHeII[MAX_NCELLS];
    double c_ciHeIS_ciHeIS[1024];
    double cs_ciHeIS_ciHeIS[MAX_NCELLS];
    double dcs_ciHeIS_ciHeIS[MAX_NCELLS];
    double c_ciHI_ciHI[1024];
    double cs_ciHI_ciHI[MAX_NCELLS];
    double dcs_ciHI_ciHI[MAX_NCELLS];
    double c_compton_comp[1024];
    double cs_compton_comp[MAX_NCELLS];
    double dcs_compton_comp[MAX_NCELLS];
    double c_gammah_gammah[1024];
    double cs_gammah_gammah[MAX_NCELLS];
    double dcs_gammah_gammah[MAX_NCELLS];
    double c_gloverabel08_gael[1024];
    double cs_gloverabel08_gael[MAX_NCELLS];
    double dcs_gloverabel08_gael[MAX_NCELLS];
    double c_gloverabel08_gaH2[1024];
    double cs_gloverabel08_gaH2[MAX_NCELLS];
    double dcs_gloverabel08_gaH2[MAX_NCELLS];
    double c_gloverabel08_gaHe[1024];
    double cs_gloverabel08_gaHe[MAX_NCELLS];
    double dcs_gloverabel08_gaHe[MAX_NCELLS];
    double c_gloverabel08_gaHI[1024];
    double cs_gloverabel08_gaHI[MAX_NCELLS];
    double dcs_gloverabel08_gaHI[MAX_NCELLS];
    double c_gloverabel08_gaHp[1024];
    double cs_gloverabel08_gaHp[MAX_NCELLS];
    double dcs_gloverabel08_gaHp[MAX_NCELLS];
    double c_gloverabel08_gphdl[1024];
    double cs_gloverabel08_gphdl[MAX_NCELLS];
    double dcs_gloverabel08_gphdl[MAX_NCELLS];
    double c_gloverabel08_gpldl[1024];
    double cs_gloverabel08_gpldl[MAX_NCELLS];
    double dcs_gloverabel08_gpldl[MAX_NCELLS];
    double c_h2formation_h2mcool[1024];
    double cs_h2formation_h2mcool[MAX_NCELLS];
    double dcs_h2formation_h2mcool[MAX_NCELLS];
    double c_h2formation_h2mheat[1024];
    double cs_h2formation_h2mheat[MAX_NCELLS];
    double dcs_h2formation_h2mheat[MAX_NCELLS];
    double c_reHeII1_reHeII1[1024];
    double cs_reHeII1_reHeII1[MAX_NCELLS];
    double dcs_reHeII1_reHeII1[MAX_NCELLS];
    double c_reHeII2_reHeII2[1024];
    double cs_reHeII2_reHeII2[MAX_NCELLS];
    double dcs_reHeII2_reHeII2[MAX_NCELLS];
    double c_reHeIII_reHeIII[1024];
    double cs_reHeIII_reHeIII[MAX_NCELLS];
    double dcs_reHeIII_reHeIII[MAX_NCELLS];
    double c_reHII_reHII[1024];
    double cs_reHII_reHII[MAX_NCELLS];
    double dcs_reHII_reHII[MAX_NCELLS];
    int bin_id[MAX_NCELLS];
    int ncells;
    /* Device-resident tables and scratch, owned by the HIP library. */
    void *device_state;
} primgen_data;

/* Table fields in struct order, for code that walks every table. */
/* Field names in solver order, as <name>_setup_data reports them. */
#define PRIMGEN_FIELD_NAMES {"H2I", "H2II", "HI", "HII", "HM", "HeI", "HeII", "HeIII", "de", "ge"}
#define PRIMGEN_TABLE_FIELDS(X) \
    X(r_k01) \
    X(r_k02) \
    X(r_k03) \
    X(r_k04) \
    X(r_k05) \
    X(r_k06) \
    X(r_k07) \
    X(r_k08) \
    X(r_k09) \
    X(r_k10) \
    X(r_k11) \
    X(r_k12) \
    X(r_k13) \
    X(r_k14) \
    X(r_k15) \
    X(r_k16) \
    X(r_k17) \
    X(r_k18) \
    X(r_k19) \
    X(r_k21) \
    X(r_k22) \
    X(r_k23) \
    X(c_brem_brem) \
    X(c_ceHeI_ceHeI) \
    X(c_ceHeII_ceHeII) \
    X(c_ceHI_ceHI) \
    X(c_ciHeI_ciHeI) \
    X(c_ciHeII_ciHeII) \
    X(c_ciHeIS_ciHeIS) \
    X(c_ciHI_ciHI) \
    X(c_compton_comp) \
    X(c_gammah_gammah) \
    X(c_gloverabel08_gael) \
    X(c_gloverabel08_gaH2) \
    X(c_gloverabel08_gaHe) \
    X(c_gloverabel08_gaHI) \
    X(c_gloverabel08_gaHp) \
    X(c_gloverabel08_gphdl) \
    X(c_gloverabel08_gpldl) \
    X(c_h2formation_h2mcool) \
    X(c_h2formation_h2mheat) \
    X(c_reHeII1_reHeII1) \
    X(c_reHeII2_reHeII2) \
    X(c_reHeIII_reHeIII) \
    X(c_reHII_reHII)
/* Values per table in struct order: the T grid, or the z grid for a redshift table. */
#define PRIMGEN_TABLE_LENGTHS {1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024}

#include <stddef.h>

#include "dengo_types.h"

#ifdef __cplusplus
extern "C" {
#endif
/* Drop-in entry points (rates_and_rate_tables.C.template:17-54, 149-274,
   278-304, 309-350, 460-599; the reference's double &dt is double *dt). */
primgen_data *primgen_setup_data(int *NumberOfFields, char ***FieldNames);
void primgen_read_rate_tables(primgen_data *data);
void primgen_read_cooling_tables(primgen_data *data);
void primgen_free_data(primgen_data *data);
/* rates_and_rate_tables.C.template:56-147: reads primgen_initial_conditions.h5,
   integrates to PRIMGEN_STOP_TIME, writes primgen_solution.h5 (the .bin forms of
   dengo_amd.tables without HDF5); 0 on success, 1 on failure. */
int primgen_main(int argc, char **argv);
double dengo_evolve_primgen(double dtf, double *dt, double z, double *input,
        double *rtol, double *atol, long long dims, primgen_data *data);
int calculate_rhs_primgen(double *input, double *rhs, int nstrip, int nchem, void *sdata);
int calculate_jacobian_primgen(double *input, double *Joutput, int nstrip, int nchem, void *sdata);
void primgen_calculate_temperature(primgen_data *data, double *input, int nstrip, int nchem);

/* Backend extensions (no counterpart in the reference).
   Allocate an empty struct; set tables from memory (tables[t * nT + i] in
   struct order, T grid log-spaced on [T0, T1]); read the table blob from an
   explicit path. */
primgen_data *primgen_alloc_data(void);
int primgen_set_tables(primgen_data *data, const double *tables, int ntables, int nT,
        double T0, double T1);
primgen_data *primgen_setup_data_from(const char *path, int *NumberOfFields, char ***FieldNames);
/* sizeof(primgen_data), for bindings that hold the struct as bytes. */
size_t primgen_data_bytes(void);
/* Integrate dims cells held in device memory (AoS [dims][NSPECIES], updated
   in place), enqueued on stream (a hipStream_t, NULL = default).  dt0 < 0
   selects dtf/1e5.  With ttot_min non-NULL the call synchronises and returns
   the smallest elapsed time and that cell's final dt. */
int dengo_primgen_evolve_device(primgen_data *data, const dengo_options *opt, double dtf,
        double dt0, double z, double *d_input, const double *d_rtol, double *d_atol,
        long long dims, const dengo_cell_outputs *outputs, double *ttot_min,
        double *dt_min, void *stream);
/* Host-array form of dengo_evolve_primgen with options and per-cell outputs. */
double dengo_primgen_evolve_host(primgen_data *data, const dengo_options *opt, double dtf,
        double *dt, double z, double *input, double *rtol, double *atol, long long dims,
        int *status, int *nsub, int *nrej, int *nsweep, double *dt_cell, double *T_cell);
/* Batched RHS f[dims][N], Jacobian J[dims][N][N] (column-major per cell, may
   be NULL), temperature T[dims] and interpolated tables rates[dims][NTABLES]
   (may be NULL) for number-density states on host memory; returns the number
   of cells whose RHS saw a negative value (their f and J are zero), or a
   negative error code. */
long long dengo_primgen_eval(primgen_data *data, const double *u, long long dims, double z,
        double *f, double *J, double *T, double *rates);
/* BE_chem_solve(calculate_rhs_primgen, calculate_jacobian_primgen, ...) as the Cython
   harness calls it (dengo/solvers/BE_chem_solve.C:79-241,
   rates_and_rate_tables_run.pyx.template:142) on host arrays: one backward-Euler
   step of nstrip cells that share dt and the Newton convergence decision; u and
   atol are scaled and unscaled in place as the reference does; the redshift is
   data->current_z.  Returns 0 (converged), 1 (not converged in 10 sweeps, or a
   failed RHS/Jacobian/NaN update), or a negative error code. */
int dengo_primgen_be_chem_solve(primgen_data *data, double *u, double dt, double *rtol,
        double *atol, int nstrip, int nchem, double *scaling);
/* ensure_electron_consistency (rates_and_rate_tables.C.template:604-630): de of
   cell i becomes the free-electron total of cells 0..i (the reference's total
   runs across the strip). */
void ensure_electron_consistency_primgen(double *input, int nstrip, int nchem);
/* temperature_from_mass_density (rates_and_rate_tables.C.template:635-668): the
   temperature of nstrip cells given as mass densities (ge in erg/g), clamped below
   at 1 K only; computed on the GPU. */
void temperature_from_mass_density_primgen(double *input, int nstrip, int nchem,
        double *strip_temperature);
/* The same two helpers under the reference's unsuffixed names
   (rates_and_rate_tables.h.template:106-107): this library holds one network. */
void ensure_electron_consistency(double *input, int nstrip, int nchem);
void temperature_from_mass_density(double *input, int nstrip, int nchem,
        double *strip_temperature);
/* Turns the intermediate-solution files of dengo_evolve_primgen and primgen_main on (1)
   or off (0); the default is PRIMGEN_WRITE_INTERMEDIATE.  While on, the host-array
   evolve runs the reference-faithful FP64 sweep and, after it, writes file k for
   every k up to the largest accepted-substep count (at most what 1 GiB of device
   records holds): each cell's state before its k-th accepted substep (number
   densities), /T of its last sweep, the root attributes time and timestep of
   the cell with the least elapsed time, and per-cell cell_time and cell_timestep
   attributes; a cell with fewer substeps repeats its last record (NaN without
   one).  Returns 0, or a negative error code. */
int dengo_primgen_set_write_intermediate(primgen_data *data, int on);
/* This network's hooks for dengo_server_add_network (include/dengo_server.h):
   a patch server in the same process can then serve it. */
const dengo_net_hooks *dengo_primgen_server_hooks(void);
/* Sets data->current_z, the redshift the strip solve uses. */
void dengo_primgen_set_current_z(primgen_data *data, double z);
/* Copies data->Ts (temperatures of the last strip RHS evaluation) for
   min(nstrip, MAX_NCELLS) cells; returns that count. */
int dengo_primgen_strip_temperatures(const primgen_data *data, double *T, int nstrip);
#ifdef __cplusplus
}
static inline double dengo_evolve_primgen(double dtf, double &dt, double z, double *input,
        double *rtol, double *atol, long long dims, primgen_data *data)
{
    return dengo_evolve_primgen(dtf, &dt, z, input, rtol, atol, dims, data);
}
#endif

#endif
