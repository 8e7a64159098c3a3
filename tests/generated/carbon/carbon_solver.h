/* Drop-in header for the carbon network (generated by dengo_amd.codegen).
   Struct fields and order follow dengo/templates/rates_and_rate_tables.h.template:30-83;
   entry points follow :85-108 with C linkage. */
#ifndef CARBON_SOLVER_H
#define CARBON_SOLVER_H

#ifndef MAX_NCELLS
#define MAX_NCELLS 1024
#endif
#define CARBON_NSPECIES 9
#define CARBON_NTABLES 19
#define CARBON_NT 1024
/* Temperature grid of the tables and <name>_main's dtf
   (rates_and_rate_tables.C.template:25-26, 110). */
#define CARBON_T_LO 10000.0
#define CARBON_T_HI 1000000000.0
#define CARBON_STOP_TIME 1000000000000.0
/* ChemicalNetwork(write_intermediate=...): dengo_evolve writes
   <name>_intermediate_%%06d.h5 after every accepted substep
   (rates_and_rate_tables.C.template:212-241). */
#define CARBON_WRITE_INTERMEDIATE 0
/* Redshift grid of the pi/ph tables (rates_and_rate_tables.C.template:31-36). */
#define CARBON_NZ 1
#define CARBON_Z_LO 0.0
#define CARBON_Z_HI 0.0

typedef struct carbon_data_ {
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
    double r_C_1_i[1024];
    double rs_C_1_i[MAX_NCELLS];
    double drs_C_1_i[MAX_NCELLS];
    double r_C_2_i[1024];
    double rs_C_2_i[MAX_NCELLS];
    double drs_C_2_i[MAX_NCELLS];
    double r_C_2_r[1024];
    double rs_C_2_r[MAX_NCELLS];
    double drs_C_2_r[MAX_NCELLS];
    double r_C_3_i[1024];
    double rs_C_3_i[MAX_NCELLS];
    double drs_C_3_i[MAX_NCELLS];
    double r_C_3_r[1024];
    double rs_C_3_r[MAX_NCELLS];
    double drs_C_3_r[MAX_NCELLS];
    double r_C_4_i[1024];
    double rs_C_4_i[MAX_NCELLS];
    double drs_C_4_i[MAX_NCELLS];
    double r_C_4_r[1024];
    double rs_C_4_r[MAX_NCELLS];
    double drs_C_4_r[MAX_NCELLS];
    double r_C_5_i[1024];
    double rs_C_5_i[MAX_NCELLS];
    double drs_C_5_i[MAX_NCELLS];
    double r_C_5_r[1024];
    double rs_C_5_r[MAX_NCELLS];
    double drs_C_5_r[MAX_NCELLS];
    double r_C_6_i[1024];
    double rs_C_6_i[MAX_NCELLS];
    double drs_C_6_i[MAX_NCELLS];
    double r_C_6_r[1024];
    double rs_C_6_r[MAX_NCELLS];
    double drs_C_6_r[MAX_NCELLS];
    double r_C_7_r[1024];
    double rs_C_7_r[MAX_NCELLS];
    double drs_C_7_r[MAX_NCELLS];
    double c_C_1_c_C_1_c[1024];
    double cs_C_1_c_C_1_c[MAX_NCELLS];
    double dcs_C_1_c_C_1_c[MAX_NCELLS];
    double c_C_2_c_C_2_c[1024];
    double cs_C_2_c_C_2_c[MAX_NCELLS];
    double dcs_C_2_c_C_2_c[MAX_NCELLS];
    double c_C_3_c_C_3_c[1024];
    double cs_C_3_c_C_3_c[MAX_NCELLS];
    double dcs_C_3_c_C_3_c[MAX_NCELLS];
    double c_C_4_c_C_4_c[1024];
    double cs_C_4_c_C_4_c[MAX_NCELLS];
    double dcs_C_4_c_C_4_c[MAX_NCELLS];
    double c_C_5_c_C_5_c[1024];
    double cs_C_5_c_C_5_c[MAX_NCELLS];
    double dcs_C_5_c_C_5_c[MAX_NCELLS];
    double c_C_6_c_C_6_c[1024];
    double cs_C_6_c_C_6_c[MAX_NCELLS];
    double dcs_C_6_c_C_6_c[MAX_NCELLS];
    double c_C_7_c_C_7_c[1024];
    double cs_C_7_c_C_7_c[MAX_NCELLS];
    double dcs_C_7_c_C_7_c[MAX_NCELLS];
    int bin_id[MAX_NCELLS];
    int ncells;
    /* Device-resident tables and scratch, owned by the HIP library. */
    void *device_state;
} carbon_data;

/* Table fields in struct order, for code that walks every table. */
/* Field names in solver order, as <name>_setup_data reports them. */
#define CARBON_FIELD_NAMES {"C_1", "C_2", "C_3", "C_4", "C_5", "C_6", "C_7", "de", "ge"}
#define CARBON_TABLE_FIELDS(X) \
    X(r_C_1_i) \
    X(r_C_2_i) \
    X(r_C_2_r) \
    X(r_C_3_i) \
    X(r_C_3_r) \
    X(r_C_4_i) \
    X(r_C_4_r) \
    X(r_C_5_i) \
    X(r_C_5_r) \
    X(r_C_6_i) \
    X(r_C_6_r) \
    X(r_C_7_r) \
    X(c_C_1_c_C_1_c) \
    X(c_C_2_c_C_2_c) \
    X(c_C_3_c_C_3_c) \
    X(c_C_4_c_C_4_c) \
    X(c_C_5_c_C_5_c) \
    X(c_C_6_c_C_6_c) \
    X(c_C_7_c_C_7_c)
/* Values per table in struct order: the T grid, or the z grid for a redshift table. */
#define CARBON_TABLE_LENGTHS {1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024, 1024}

#include <stddef.h>

#include "dengo_types.h"

#ifdef __cplusplus
extern "C" {
#endif
/* Drop-in entry points (rates_and_rate_tables.C.template:17-54, 149-274,
   278-304, 309-350, 460-599; the reference's double &dt is double *dt). */
carbon_data *carbon_setup_data(int *NumberOfFields, char ***FieldNames);
void carbon_read_rate_tables(carbon_data *data);
void carbon_read_cooling_tables(carbon_data *data);
void carbon_free_data(carbon_data *data);
/* rates_and_rate_tables.C.template:56-147: reads carbon_initial_conditions.h5,
   integrates to CARBON_STOP_TIME, writes carbon_solution.h5 (the .bin forms of
   dengo_amd.tables without HDF5); 0 on success, 1 on failure. */
int carbon_main(int argc, char **argv);
double dengo_evolve_carbon(double dtf, double *dt, double z, double *input,
        double *rtol, double *atol, long long dims, carbon_data *data);
int calculate_rhs_carbon(double *input, double *rhs, int nstrip, int nchem, void *sdata);
int calculate_jacobian_carbon(double *input, double *Joutput, int nstrip, int nchem, void *sdata);
void carbon_calculate_temperature(carbon_data *data, double *input, int nstrip, int nchem);

/* Backend extensions (no counterpart in the reference).
   Allocate an empty struct; set tables from memory (tables[t * nT + i] in
   struct order, T grid log-spaced on [T0, T1]); read the table blob from an
   explicit path. */
carbon_data *carbon_alloc_data(void);
int carbon_set_tables(carbon_data *data, const double *tables, int ntables, int nT,
        double T0, double T1);
carbon_data *carbon_setup_data_from(const char *path, int *NumberOfFields, char ***FieldNames);
/* sizeof(carbon_data), for bindings that hold the struct as bytes. */
size_t carbon_data_bytes(void);
/* Integrate dims cells held in device memory (AoS [dims][NSPECIES], updated
   in place), enqueued on stream (a hipStream_t, NULL = default).  dt0 < 0
   selects dtf/1e5.  With ttot_min non-NULL the call synchronises and returns
   the smallest elapsed time and that cell's final dt. */
int dengo_carbon_evolve_device(carbon_data *data, const dengo_options *opt, double dtf,
        double dt0, double z, double *d_input, const double *d_rtol, double *d_atol,
        long long dims, const dengo_cell_outputs *outputs, double *ttot_min,
        double *dt_min, void *stream);
/* Host-array form of dengo_evolve_carbon with options and per-cell outputs. */
double dengo_carbon_evolve_host(carbon_data *data, const dengo_options *opt, double dtf,
        double *dt, double z, double *input, double *rtol, double *atol, long long dims,
        int *status, int *nsub, int *nrej, int *nsweep, double *dt_cell, double *T_cell);
/* Batched RHS f[dims][N], Jacobian J[dims][N][N] (column-major per cell, may
   be NULL), temperature T[dims] and interpolated tables rates[dims][NTABLES]
   (may be NULL) for number-density states on host memory; returns the number
   of cells whose RHS saw a negative value (their f and J are zero), or a
   negative error code. */
long long dengo_carbon_eval(carbon_data *data, const double *u, long long dims, double z,
        double *f, double *J, double *T, double *rates);
/* BE_chem_solve(calculate_rhs_carbon, calculate_jacobian_carbon, ...) as the Cython
   harness calls it (dengo/solvers/BE_chem_solve.C:79-241,
   rates_and_rate_tables_run.pyx.template:142) on host arrays: one backward-Euler
   step of nstrip cells that share dt and the Newton convergence decision; u and
   atol are scaled and unscaled in place as the reference does; the redshift is
   data->current_z.  Returns 0 (converged), 1 (not converged in 10 sweeps, or a
   failed RHS/Jacobian/NaN update), or a negative error code. */
int dengo_carbon_be_chem_solve(carbon_data *data, double *u, double dt, double *rtol,
        double *atol, int nstrip, int nchem, double *scaling);
/* ensure_electron_consistency (rates_and_rate_tables.C.template:604-630): de of
   cell i becomes the free-electron total of cells 0..i (the reference's total
   runs across the strip). */
void ensure_electron_consistency_carbon(double *input, int nstrip, int nchem);
/* temperature_from_mass_density (rates_and_rate_tables.C.template:635-668): the
   temperature of nstrip cells given as mass densities (ge in erg/g), clamped below
   at 1 K only; computed on the GPU. */
void temperature_from_mass_density_carbon(double *input, int nstrip, int nchem,
        double *strip_temperature);
/* The same two helpers under the reference's unsuffixed names
   (rates_and_rate_tables.h.template:106-107): this library holds one network. */
void ensure_electron_consistency(double *input, int nstrip, int nchem);
void temperature_from_mass_density(double *input, int nstrip, int nchem,
        double *strip_temperature);
/* Turns the intermediate-solution files of dengo_evolve_carbon and carbon_main on (1)
   or off (0); the default is CARBON_WRITE_INTERMEDIATE.  While on, the host-array
   evolve runs the reference-faithful FP64 sweep and, after it, writes file k for
   every k up to the largest accepted-substep count (at most what 1 GiB of device
   records holds): each cell's state before its k-th accepted substep (number
   densities), /T of its last sweep, the root attributes time and timestep of
   the cell with the least elapsed time, and per-cell cell_time and cell_timestep
   attributes; a cell with fewer substeps repeats its last record (NaN without
   one).  Returns 0, or a negative error code. */
int dengo_carbon_set_write_intermediate(carbon_data *data, int on);
/* This network's hooks for dengo_server_add_network (include/dengo_server.h):
   a patch server in the same process can then serve it. */
const dengo_net_hooks *dengo_carbon_server_hooks(void);
/* Sets data->current_z, the redshift the strip solve uses. */
void dengo_carbon_set_current_z(carbon_data *data, double z);
/* Copies data->Ts (temperatures of the last strip RHS evaluation) for
   min(nstrip, MAX_NCELLS) cells; returns that count. */
int dengo_carbon_strip_temperatures(const carbon_data *data, double *T, int nstrip);
#ifdef __cplusplus
}
static inline double dengo_evolve_carbon(double dtf, double &dt, double z, double *input,
        double *rtol, double *atol, long long dims, carbon_data *data)
{
    return dengo_evolve_carbon(dtf, &dt, z, input, rtol, atol, dims, data);
}
#endif

#endif
