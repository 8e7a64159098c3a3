/* Drop-in header for the oxydata network (generated by dengo_amd.codegen).
   Struct fields and order follow dengo/templates/rates_and_rate_tables.h.template:30-83;
   entry points follow :85-108 with C linkage. */
#ifndef OXYDATA_SOLVER_H
#define OXYDATA_SOLVER_H

#ifndef MAX_NCELLS
#define MAX_NCELLS 1024
#endif
#define OXYDATA_NSPECIES 11
#define OXYDATA_NTABLES 41
#define OXYDATA_NT 1024
/* Temperature grid of the tables and <name>_main's dtf
   (rates_and_rate_tables.C.template:25-26, 110). */
#define OXYDATA_T_LO 100.0
#define OXYDATA_T_HI 1000000000.0
#define OXYDATA_STOP_TIME 1000000000000.0
/* ChemicalNetwork(write_intermediate=...): dengo_evolve writes
   <name>_intermediate_%%06d.h5 after every accepted substep
   (rates_and_rate_tables.C.template:212-241). */
#define OXYDATA_WRITE_INTERMEDIATE 0
/* Redshift grid of the pi/ph tables (rates_and_rate_tables.C.template:31-36). */
#define OXYDATA_NZ 100
#define OXYDATA_Z_LO 0.0
#define OXYDATA_Z_HI 9.0

typedef struct oxydata_data_ {
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
    double r_O_1_i[1024];
    double rs_O_1_i[MAX_NCELLS];
    double drs_O_1_i[MAX_NCELLS];
    double r_O_1_pi[100];
    double rs_O_1_pi[MAX_NCELLS];
    double drs_O_1_pi[MAX_NCELLS];
    double r_O_2_i[1024];
    double rs_O_2_i[MAX_NCELLS];
    double drs_O_2_i[MAX_NCELLS];
    double r_O_2_pi[100];
    double rs_O_2_pi[MAX_NCELLS];
    double drs_O_2_pi[MAX_NCELLS];
    double r_O_2_r[1024];
    double rs_O_2_r[MAX_NCELLS];
    double drs_O_2_r[MAX_NCELLS];
    double r_O_3_i[1024];
    double rs_O_3_i[MAX_NCELLS];
    double drs_O_3_i[MAX_NCELLS];
    double r_O_3_pi[100];
    double rs_O_3_pi[MAX_NCELLS];
    double drs_O_3_pi[MAX_NCELLS];
    double r_O_3_r[1024];
    double rs_O_3_r[MAX_NCELLS];
    double drs_O_3_r[MAX_NCELLS];
    double r_O_4_i[1024];
    double rs_O_4_i[MAX_NCELLS];
    double drs_O_4_i[MAX_NCELLS];
    double r_O_4_pi[100];
    double rs_O_4_pi[MAX_NCELLS];
    double drs_O_4_pi[MAX_NCELLS];
    double r_O_4_r[1024];
    double rs_O_4_r[MAX_NCELLS];
    double drs_O_4_r[MAX_NCELLS];
    double r_O_5_i[1024];
    double rs_O_5_i[MAX_NCELLS];
    double drs_O_5_i[MAX_NCELLS];
    double r_O_5_pi[100];
    double rs_O_5_pi[MAX_NCELLS];
    double drs_O_5_pi[MAX_NCELLS];
    double r_O_5_r[1024];
    double rs_O_5_r[MAX_NCELLS];
    double drs_O_5_r[MAX_NCELLS];
    double r_O_6_i[1024];
    double rs_O_6_i[MAX_NCELLS];
    double drs_O_6_i[MAX_NCELLS];
    double r_O_6_pi[100];
    double rs_O_6_pi[MAX_NCELLS];
    double drs_O_6_pi[MAX_NCELLS];
    double r_O_6_r[1024];
    double rs_O_6_r[MAX_NCELLS];
    double drs_O_6_r[MAX_NCELLS];
    double r_O_7_i[1024];
    double rs_O_7_i[MAX_NCELLS];
    double drs_O_7_i[MAX_NCELLS];
    double r_O_7_pi[100];
    double rs_O_7_pi[MAX_NCELLS];
    double drs_O_7_pi[MAX_NCELLS];
    double r_O_7_r[1024];
    double rs_O_7_r[MAX_NCELLS];
    double drs_O_7_r[MAX_NCELLS];
    double r_O_8_i[1024];
    double rs_O_8_i[MAX_NCELLS];
    double drs_O_8_i[MAX_NCELLS];
    double r_O_8_pi[100];
    double rs_O_8_pi[MAX_NCELLS];
    double drs_O_8_pi[MAX_NCELLS];
    double r_O_8_r[1024];
    double rs_O_8_r[MAX_NCELLS];
    double drs_O_8_r[MAX_NCELLS];
    double r_O_9_r[1024];
    double rs_O_9_r[MAX_NCELLS];
    double drs_O_9_r[MAX_NCELLS];
    double c_O_1_c_O_1_c[1024];
    double cs_O_1_c_O_1_c[MAX_NCELLS];
    double dcs_O_1_c_O_1_c[MAX_NCELLS];
    double c_O_1_ph_O_1_ph[100];
    double cs_O_1_ph_O_1_ph[MAX_NCELLS];
    double dcs_O_1_ph_O_1_ph[MAX_NCELLS];
    double c_O_2_c_O_2_c[1024];
    double cs_O_2_c_O_2_c[MAX_NCELLS];
    double dcs_O_2_c_O_2_c[MAX_NCELLS];
    double c_O_2_ph_O_2_ph[100];
    double cs_O_2_ph_O_2_ph[MAX_NCELLS];
    double dcs_O_2_ph_O_2_ph[MAX_NCELLS];
    double c_O_3_c_O_3_c[1024];
    double cs_O_3_c_O_3_c[MAX_NCELLS];
    double dcs_O_3_c_O_3_c[MAX_NCELLS];
    double c_O_3_ph_O_3_ph[100];
    double cs_O_3_ph_O_3_ph[MAX_NCELLS];
    double dcs_O_3_ph_O_3_ph[MAX_NCELLS];
    double c_O_4_c_O_4_c[1024];
    double cs_O_4_c_O_4_c[MAX_NCELLS];
    double dcs_O_4_c_O_4_c[MAX_NCELLS];
    double c_O_4_ph_O_4_ph[100];
    double cs_O_4_ph_O_4_ph[MAX_NCELLS];
    double dcs_O_4_ph_O_4_ph[MAX_NCELLS];
    double c_O_5_c_O_5_c[1024];
    double cs_O_5_c_O_5_c[MAX_NCELLS];
    double dcs_O_5_c_O_5_c[MAX_NCELLS];
    double c_O_5_ph_O_5_ph[100];
    double cs_O_5_ph_O_5_ph[MAX_NCELLS];
    double dcs_O_5_ph_O_5_ph[MAX_NCELLS];
    double c_O_6_c_O_6_c[1024];
    double cs_O_6_c_O_6_c[MAX_NCELLS];
    double dcs_O_6_c_O_6_c[MAX_NCELLS];
    double c_O_6_ph_O_6_ph[100];
    double cs_O_6_ph_O_6_ph[MAX_NCELLS];
    double dcs_O_6_ph_O_6_ph[MAX_NCELLS];
    double c_O_7_c_O_7_c[1024];
    double cs_O_7_c_O_7_c[MAX_NCELLS];
    double dcs_O_7_c_O_7_c[MAX_NCELLS];
    double c_O_7_ph_O_7_ph[100];
    double cs_O_7_ph_O_7_ph[MAX_NCELLS];
    double dcs_O_7_ph_O_7_ph[MAX_NCELLS];
    double c_O_8_c_O_8_c[1024];
    double cs_O_8_c_O_8_c[MAX_NCELLS];
    double dcs_O_8_c_O_8_c[MAX_NCELLS];
    double c_O_8_ph_O_8_ph[100];
    double cs_O_8_ph_O_8_ph[MAX_NCELLS];
    double dcs_O_8_ph_O_8_ph[MAX_NCELLS];
    double c_O_9_c_O_9_c[1024];
    double cs_O_9_c_O_9_c[MAX_NCELLS];
    double dcs_O_9_c_O_9_c[MAX_NCELLS];
    int bin_id[MAX_NCELLS];
    int ncells;
    /* Device-resident tables and scratch, owned by the HIP library. */
    void *device_state;
} oxydata_data;

/* Table fields in struct order, for code that walks every table. */
/* Field names in solver order, as <name>_setup_data reports them. */
#define OXYDATA_FIELD_NAMES {"O_1", "O_2", "O_3", "O_4", "O_5", "O_6", "O_7", "O_8", "O_9", "de", "ge"}
#define OXYDATA_TABLE_FIELDS(X) \
    X(r_O_1_i) \
    X(r_O_1_pi) \
    X(r_O_2_i) \
    X(r_O_2_pi) \
    X(r_O_2_r) \
    X(r_O_3_i) \
    X(r_O_3_pi) \
    X(r_O_3_r) \
    X(r_O_4_i) \
    X(r_O_4_pi) \
    X(r_O_4_r) \
    X(r_O_5_i) \
    X(r_O_5_pi) \
    X(r_O_5_r) \
    X(r_O_6_i) \
    X(r_O_6_pi) \
    X(r_O_6_r) \
    X(r_O_7_i) \
    X(r_O_7_pi) \
    X(r_O_7_r) \
    X(r_O_8_i) \
    X(r_O_8_pi) \
    X(r_O_8_r) \
    X(r_O_9_r) \
    X(c_O_1_c_O_1_c) \
    X(c_O_1_ph_O_1_ph) \
    X(c_O_2_c_O_2_c) \
    X(c_O_2_ph_O_2_ph) \
    X(c_O_3_c_O_3_c) \
    X(c_O_3_ph_O_3_ph) \
    X(c_O_4_c_O_4_c) \
    X(c_O_4_ph_O_4_ph) \
    X(c_O_5_c_O_5_c) \
    X(c_O_5_ph_O_5_ph) \
    X(c_O_6_c_O_6_c) \
    X(c_O_6_ph_O_6_ph) \
    X(c_O_7_c_O_7_c) \
    X(c_O_7_ph_O_7_ph) \
    X(c_O_8_c_O_8_c) \
    X(c_O_8_ph_O_8_ph) \
    X(c_O_9_c_O_9_c)
/* Values per table in struct order: the T grid, or the z grid for a redshift table. */
#define OXYDATA_TABLE_LENGTHS {1024, 100, 1024, 100, 1024, 1024, 100, 1024, 1024, 100, 1024, 1024, 100, 1024, 1024, 100, 1024, 1024, 100, 1024, 1024, 100, 1024, 1024, 1024, 100, 1024, 100, 1024, 100, 1024, 100, 1024, 100, 1024, 100, 1024, 100, 1024, 100, 1024}

#include <stddef.h>

#include "dengo_types.h"

#ifdef __cplusplus
extern "C" {
#endif
/* Drop-in entry points (rates_and_rate_tables.C.template:17-54, 149-274,
   278-304, 309-350, 460-599; the reference's double &dt is double *dt). */
oxydata_data *oxydata_setup_data(int *NumberOfFields, char ***FieldNames);
void oxydata_read_rate_tables(oxydata_data *data);
void oxydata_read_cooling_tables(oxydata_data *data);
void oxydata_free_data(oxydata_data *data);
/* rates_and_rate_tables.C.template:56-147: reads oxydata_initial_conditions.h5,
   integrates to OXYDATA_STOP_TIME, writes oxydata_solution.h5 (the .bin forms of
   dengo_amd.tables without HDF5); 0 on success, 1 on failure. */
int oxydata_main(int argc, char **argv);
double dengo_evolve_oxydata(double dtf, double *dt, double z, double *input,
        double *rtol, double *atol, long long dims, oxydata_data *data);
int calculate_rhs_oxydata(double *input, double *rhs, int nstrip, int nchem, void *sdata);
int calculate_jacobian_oxydata(double *input, double *Joutput, int nstrip, int nchem, void *sdata);
void oxydata_calculate_temperature(oxydata_data *data, double *input, int nstrip, int nchem);

/* Backend extensions (no counterpart in the reference).
   Allocate an empty struct; set tables from memory (tables[t * nT + i] in
   struct order, T grid log-spaced on [T0, T1]); read the table blob from an
   explicit path. */
oxydata_data *oxydata_alloc_data(void);
int oxydata_set_tables(oxydata_data *data, const double *tables, int ntables, int nT,
        double T0, double T1);
oxydata_data *oxydata_setup_data_from(const char *path, int *NumberOfFields, char ***FieldNames);
/* sizeof(oxydata_data), for bindings that hold the struct as bytes. */
size_t oxydata_data_bytes(void);
/* Integrate dims cells held in device memory (AoS [dims][NSPECIES], updated
   in place), enqueued on stream (a hipStream_t, NULL = default).  dt0 < 0
   selects dtf/1e5.  With ttot_min non-NULL the call synchronises and returns
   the smallest elapsed time and that cell's final dt. */
int dengo_oxydata_evolve_device(oxydata_data *data, const dengo_options *opt, double dtf,
        double dt0, double z, double *d_input, const double *d_rtol, double *d_atol,
        long long dims, const dengo_cell_outputs *outputs, double *ttot_min,
        double *dt_min, void *stream);
/* Host-array form of dengo_evolve_oxydata with options and per-cell outputs. */
double dengo_oxydata_evolve_host(oxydata_data *data, const dengo_options *opt, double dtf,
        double *dt, double z, double *input, double *rtol, double *atol, long long dims,
        int *status, int *nsub, int *nrej, int *nsweep, double *dt_cell, double *T_cell);
/* Batched RHS f[dims][N], Jacobian J[dims][N][N] (column-major per cell, may
   be NULL), temperature T[dims] and interpolated tables rates[dims][NTABLES]
   (may be NULL) for number-density states on host memory; returns the number
   of cells whose RHS saw a negative value (their f and J are zero), or a
   negative error code. */
long long dengo_oxydata_eval(oxydata_data *data, const double *u, long long dims, double z,
        double *f, double *J, double *T, double *rates);
/* BE_chem_solve(calculate_rhs_oxydata, calculate_jacobian_oxydata, ...) as the Cython
   harness calls it (dengo/solvers/BE_chem_solve.C:79-241,
   rates_and_rate_tables_run.pyx.template:142) on host arrays: one backward-Euler
   step of nstrip cells that share dt and the Newton convergence decision; u and
   atol are scaled and unscaled in place as the reference does; the redshift is
   data->current_z.  Returns 0 (converged), 1 (not converged in 10 sweeps, or a
   failed RHS/Jacobian/NaN update), or a negative error code. */
int dengo_oxydata_be_chem_solve(oxydata_data *data, double *u, double dt, double *rtol,
        double *atol, int nstrip, int nchem, double *scaling);
/* ensure_electron_consistency (rates_and_rate_tables.C.template:604-630): de of
   cell i becomes the free-electron total of cells 0..i (the reference's total
   runs across the strip). */
void ensure_electron_consistency_oxydata(double *input, int nstrip, int nchem);
/* temperature_from_mass_density (rates_and_rate_tables.C.template:635-668): the
   temperature of nstrip cells given as mass densities (ge in erg/g), clamped below
   at 1 K only; computed on the GPU. */
void temperature_from_mass_density_oxydata(double *input, int nstrip, int nchem,
        double *strip_temperature);
/* The same two helpers under the reference's unsuffixed names
   (rates_and_rate_tables.h.template:106-107): this library holds one network. */
void ensure_electron_consistency(double *input, int nstrip, int nchem);
void temperature_from_mass_density(double *input, int nstrip, int nchem,
        double *strip_temperature);
/* Turns the intermediate-solution files of dengo_evolve_oxydata and oxydata_main on (1)
   or off (0); the default is OXYDATA_WRITE_INTERMEDIATE.  While on, the host-array
   evolve runs the reference-faithful FP64 sweep and, after it, writes file k for
   every k up to the largest accepted-substep count (at most what 1 GiB of device
   records holds): each cell's state before its k-th accepted substep (number
   densities), /T of its last sweep, the root attributes time and timestep of
   the cell with the least elapsed time, and per-cell cell_time and cell_timestep
   attributes; a cell with fewer substeps repeats its last record (NaN without
   one).  Returns 0, or a negative error code. */
int dengo_oxydata_set_write_intermediate(oxydata_data *data, int on);
/* This network's hooks for dengo_server_add_network (include/dengo_server.h):
   a patch server in the same process can then serve it. */
const dengo_net_hooks *dengo_oxydata_server_hooks(void);
/* Sets data->current_z, the redshift the strip solve uses. */
void dengo_oxydata_set_current_z(oxydata_data *data, double z);
/* Copies data->Ts (temperatures of the last strip RHS evaluation) for
   min(nstrip, MAX_NCELLS) cells; returns that count. */
int dengo_oxydata_strip_temperatures(const oxydata_data *data, double *T, int nstrip);
#ifdef __cplusplus
}
static inline double dengo_evolve_oxydata(double dtf, double &dt, double z, double *input,
        double *rtol, double *atol, long long dims, oxydata_data *data)
{
    return dengo_evolve_oxydata(dtf, &dt, z, input, rtol, atol, dims, data);
}
#endif

#endif
