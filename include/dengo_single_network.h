/*
 * The reference's unsuffixed helpers (rates_and_rate_tables.h.template:106-107).
 *
 * Exported by every single-network solver library that
 * ChemicalNetwork.write_solver emits and dengo_amd.build.build_solver builds
 * (each also exports the suffixed forms ensure_electron_consistency_<n> and
 * temperature_from_mass_density_<n>).  libdengo_hip.so holds several networks
 * and exports only the suffixed forms.
 */
#ifndef DENGO_SINGLE_NETWORK_H
#define DENGO_SINGLE_NETWORK_H

#ifdef __cplusplus
extern "C" {
#endif

/* ensure_electron_consistency (rates_and_rate_tables.C.template:604-630): de
 * of cell i becomes the free-electron total of cells 0..i of input[nstrip][nchem]
 * (the reference's running total is not reset per cell). */
void ensure_electron_consistency(double *input, int nstrip, int nchem);

/* temperature_from_mass_density (rates_and_rate_tables.C.template:635-668):
 * strip_temperature[i] for cells of mass densities (ge in erg/g), clamped below
 * at 1 K and not above. */
void temperature_from_mass_density(double *input, int nstrip, int nchem,
                                   double *strip_temperature);

#ifdef __cplusplus
}
#endif
#endif
