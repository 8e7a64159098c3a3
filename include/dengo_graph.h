/*
 * Captured launches: one integration's device work recorded once as a HIP
 * graph and replayed on any stream.
 *
 * Backend extension with no counterpart in the reference.  A hydro code that
 * calls the chemistry every step on the same device buffers (d_input, d_rtol,
 * d_atol and the per-cell outputs, as for <n>_evolve_device) pays the launch
 * sequence -- the queue-counter reset and the persistent kernel -- once at
 * capture instead of every step.  Each replay integrates whatever the
 * buffers hold at that moment, in place, with the options, dtf, dt0 and z of
 * the capture.  The graph owns its queue counter and, when opt->order is
 * given, its own gather/scatter buffers for that order (sized at creation),
 * so later calls on the same data struct and other graphs do not disturb
 * it: replays of one graph must not overlap (replays of different graphs
 * may).  The captured sequence contains kernels only (no memset or copy
 * nodes).
 */
#ifndef DENGO_GRAPH_H
#define DENGO_GRAPH_H

#include "dengo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dengo_graph dengo_graph;

/* network: enum dengo_network (dengo_server.h); data: that network's data
 * struct with tables set; d_T: per-cell temperatures of the isothermal
 * synthetic network (NULL otherwise).  Returns NULL on failure
 * (dengo_last_error()). */
dengo_graph *dengo_graph_create(int network, void *data, const dengo_options *opt, double dtf,
        double dt0, double z, const double *d_T, double *d_input, const double *d_rtol,
        double *d_atol, long long dims, const dengo_cell_outputs *outputs);

/* Enqueues one replay on stream (a hipStream_t, NULL = default). */
int dengo_graph_launch(dengo_graph *g, void *stream);

void dengo_graph_destroy(dengo_graph *g);

#ifdef __cplusplus
}
#endif
#endif
