/*
 * C ABI of the MI355X backend for Dengo's backward-Euler chemistry solver.
 *
 * Each built-in network <n> (primordial, oxygen) exports, with C linkage:
 *   <n>_setup_data        <- rates_and_rate_tables.C.template:17-54
 *   <n>_read_rate_tables / <n>_read_cooling_tables  <- :278-304
 *   dengo_evolve_<n>      <- :149-274 (double &dt becomes double *dt)
 *   calculate_rhs_<n>     <- :460-522
 *   calculate_jacobian_<n> <- :527-599
 *   <n>_calculate_temperature <- :309-350
 *   ensure_electron_consistency_<n>, temperature_from_mass_density_<n> <- :604-668
 *   <n>_main              <- :56-147
 * plus the backend extensions declared in <n>_solver.h.
 *
 * Semantics of the drop-in: every cell is integrated as its own strip of one
 * cell (dengo_evolve with nstrip = 1): it keeps its own dt, its own Newton
 * convergence test and its own electron total.  The return value is the
 * smallest elapsed time over cells (< dtf means some cell did not finish, the
 * reference's only failure signal) and *dt is that cell's final substep.
 * <n>_setup_data reads the reference's <n>_tables.h5 from the working
 * directory (libhdf5_hl opened at run time); without that file or without
 * HDF5 it reads <n>_tables.bin (a 64-byte header "DENGOTB1", int32 ntables,
 * int32 nT, double T0, T1, z0, z1, then FP64 [ntables][nT]).
 */
#ifndef DENGO_HIP_H
#define DENGO_HIP_H

#include "dengo_types.h"
#include "primordial_solver.h"
#include "oxygen_solver.h"

#endif
