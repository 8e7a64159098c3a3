/*
 * Allocation-free integration of the primordial network.
 *
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc -- this is dengo_primordial_evolve_device with the same
 * arguments, kernels and results, whose scratch (the persistent schedule's
 * work counter, the smallest-elapsed-time reduction and, when ttot_min is
 * asked for without outputs->ttot and outputs->dt, per-cell elapsed times and
 * final substeps) is a fixed static device array of the library instead of
 * memory it allocates.  It calls no hipMalloc or hipFree.
 *
 * Limits and rules:
 *  - the data struct's tables must already be resident on the current device
 *    (any earlier dengo_primordial_evolve_device, _evolve_host or _eval call
 *    on that device uploads them); otherwise the call fails;
 *  - with ttot_min non-NULL and outputs->ttot or outputs->dt NULL, at most
 *    DENGO_NOALLOC_MAX_CELLS cells (larger calls fail with an error); no limit
 *    otherwise;
 *  - opt->order must be NULL (an order needs gather buffers);
 *  - each GPU holds one copy of the array (about 32 MiB) until the process
 *    ends, shared by every call on that GPU: calls that overlap on one GPU
 *    (on different streams or threads) race and give wrong results;
 *  - with ttot_min NULL the call only enqueues work on stream (it can be
 *    captured into a graph); with ttot_min it synchronises the stream.
 */
#ifndef DENGO_NOALLOC_H
#define DENGO_NOALLOC_H

#include "dengo_types.h"
#include "primordial_solver.h"

#define DENGO_NOALLOC_MAX_CELLS 2097152LL

#ifdef __cplusplus
extern "C" {
#endif

int dengo_primordial_evolve_device_noalloc(primordial_data *data, const dengo_options *opt,
        double dtf, double dt0, double z, double *d_input, const double *d_rtol,
        double *d_atol, long long dims, const dengo_cell_outputs *outputs, double *ttot_min,
        double *dt_min, void *stream);

#ifdef __cplusplus
}
#endif
#endif
