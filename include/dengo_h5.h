/*
 * HDF5 files in the reference's layouts, with libhdf5 loaded at run time.
 *
 * The reference writes its rate tables to <n>_tables.h5, one FP64 dataset of
 * nT values per table named /<reaction> or /<action>_<table>
 * (dengo/chemical_network.py:310-318), and its initial conditions to
 * <n>_initial_conditions.h5, one dataset per species (:320-328); the emitted C
 * reads them with H5LTread_dataset_double (rates_and_rate_tables.C.template:
 * 63-107, 278-304) and <n>_main writes <n>_solution.h5 (:116-141).  These
 * entry points read and write the same layouts.  libhdf5_hl (1.10 or later) is
 * opened with dlopen on first use: $DENGO_HDF5_HL if set, else the loader's
 * search path, else the usual install locations.  Without it every call
 * fails with a message and dengo_hdf5_available() returns 0; the table and
 * initial-condition readers then fall back to this backend's binary files.
 *
 * Backend extension: the reference links HDF5 directly and has no such API.
 */
#ifndef DENGO_H5_H
#define DENGO_H5_H

#ifdef __cplusplus
extern "C" {
#endif

typedef long long dengo_h5_file;   /* an open file, or a negative code */

/* 1 when libhdf5_hl could be loaded, else 0 (reason in dengo_last_error()). */
int dengo_hdf5_available(void);

/* mode 0: read only; 1: create (truncate); 2: read-write. */
dengo_h5_file dengo_h5_open(const char *path, int mode);
int dengo_h5_close(dengo_h5_file f);

/* Number of elements of a dataset (all dimensions), or a negative code when
 * it does not exist. */
long long dengo_h5_length(dengo_h5_file f, const char *dataset);

/* Reads a dataset of exactly n elements as FP64. */
int dengo_h5_read(dengo_h5_file f, const char *dataset, double *out, long long n);

/* Writes a one-dimensional FP64 dataset of n elements. */
int dengo_h5_write(dengo_h5_file f, const char *dataset, const double *data, long long n);

/* A scalar FP64 attribute of the root group (<n>_solution.h5's time, timestep). */
int dengo_h5_set_attribute(dengo_h5_file f, const char *name, double value);
int dengo_h5_get_attribute(dengo_h5_file f, const char *name, double *value);
/* FP64 root attributes of n values (an extension: the per-cell times of the
   intermediate-solution files), the number of values an attribute holds, and
   reading exactly n of them. */
int dengo_h5_set_attribute_array(dengo_h5_file f, const char *name, const double *values,
                                 long long n);
long long dengo_h5_attribute_length(dengo_h5_file f, const char *name);
int dengo_h5_get_attribute_array(dengo_h5_file f, const char *name, double *values, long long n);

/* 1 when the file starts with the HDF5 signature, 0 otherwise. */
int dengo_h5_is_hdf5(const char *path);

#ifdef __cplusplus
}
#endif
#endif
