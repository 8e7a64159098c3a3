/*
 * Patch server: continuous batching of chemistry solves on one GPU.
 *
 * Backend extension with no counterpart in the reference.  The reference's
 * caller (a hydro code, or <n>_main in rates_and_rate_tables.C.template:
 * 56-147) calls dengo_evolve_<n> once per patch and waits for it.  The server
 * takes patches of mixed size and network as they arrive and keeps several
 * of them on the GPU at once:
 *
 *   - a request is admitted when one of the server's slots (a HIP stream with
 *     its own work counter) is free and its state fits in the server's HBM
 *     arena (a fixed device budget, sub-allocated first fit); requests that
 *     do not fit wait in arrival order, and a later, smaller request may pass
 *     one that waits only a bounded number of times;
 *   - an admitted patch is copied into the arena, integrated by the same
 *     kernels as <n>_evolve_device (the persistent schedule, with its grid
 *     capped so that patches in flight share the GPU), and copied back into
 *     the caller's arrays; its arena block and slot are released as soon as
 *     its copy-back completes, and the next waiting request is admitted then;
 *   - a worker thread of the server does all admission, launching and
 *     completion tracking; submit only queues the request.
 *
 * Each patch is integrated exactly as by <n>_evolve_device with the same
 * options (cells are independent), so its result does not depend on what
 * else is in flight.  Host arrays passed to submit must stay valid until the
 * request completes; pinned memory (dengo_host_alloc) makes the copies
 * asynchronous.  All times are seconds on the server's steady clock, or
 * milliseconds on the device timeline that starts at the server's creation.
 */
#ifndef DENGO_SERVER_H
#define DENGO_SERVER_H

#include <stddef.h>

#include "dengo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Networks a server holds from its creation (one data struct each, any may be
 * NULL); dengo_server_add_network adds others (ids 3, 4, ..., at most 16 in all). */
enum dengo_network { DENGO_NET_PRIMORDIAL = 0, DENGO_NET_OXYGEN = 1, DENGO_NET_SYNTHETIC = 2 };

typedef struct dengo_server dengo_server;

typedef struct dengo_server_config {
    size_t hbm_budget;    /* bytes of the device arena for patches in flight */
    int slots;            /* patches in flight at once (streams), >= 1 */
    int grid_share;       /* a patch's persistent grid is at most 1/grid_share
                             of the grid that fills the GPU (0: slots) */
    int max_bypass;       /* times a waiting request may be passed by later
                             ones that fit (0: 8) */
    int precision[3];     /* per network (enum dengo_network): the Newton
                             sweep's precision, or -1 for the one given to
                             dengo_server_create */
    int grid_shares[3];   /* per network: grid_share for its patches, or 0 for
                             grid_share (a network whose patches are long may
                             take a larger share of the GPU) */
    int dense_order;      /* 1 (default): the dense network's kernel hands out
                             each patch's cells densest first (a cell's Newton
                             work grows with its density); 0: in index order.
                             Results do not depend on it. */
} dengo_server_config;

/* Completion record of one request. */
typedef struct dengo_patch_result {
    long long ticket;
    int network;
    int error;            /* 0, or a negative code (message in `message`) */
    long long ncells;
    int slot;
    int inflight_at_admit;/* patches in flight, this one included, when it was admitted */
    int max_inflight;     /* most patches in flight at once while it was */
    int status_counts[3]; /* cells by final status (0 reached dtf, 1 gave up, 2 capped) */
    double t_submit, t_admit, t_done;   /* server steady clock, s */
    double latency_ms;    /* t_done - t_submit */
    double queue_ms;      /* t_admit - t_submit */
    double dev_start_ms, dev_end_ms;    /* kernel start/end on the device timeline */
    char message[160];
} dengo_patch_result;

typedef struct dengo_server_stats {
    long long submitted, completed, failed, waiting, inflight;
    long long cells_completed;
    size_t arena_bytes, arena_used, arena_peak;
    int slots, max_inflight_seen;
    double first_submit, last_done;      /* server steady clock, s */
} dengo_server_stats;

/* Defaults: 8 GiB arena, 4 slots, grid_share = slots for every network,
 * max_bypass 8, every network at the precision given to dengo_server_create. */
void dengo_server_default_config(dengo_server_config *cfg);

/* Creates a server on the current device for the given data structs (tables
 * set; NULL for a network it will not serve; all three may be NULL when the
 * networks are added afterwards).  precision selects the Newton
 * sweep of each network's patches unless cfg->precision overrides it for that
 * network (enum dengo_precision; the dense synthetic network supports
 * DENGO_FP64 and DENGO_FP64_FAST, and DENGO_MIXED falls back to
 * DENGO_FP64_FAST there).  Returns NULL on failure (dengo_last_error()). */
dengo_server *dengo_server_create(const dengo_server_config *cfg, void *primordial,
        void *oxygen, void *synthetic, int precision);

/* Queues one patch of ncells cells of network `network`: state/rtol/atol are
 * host arrays [ncells][N] (state and atol are updated in place at
 * completion, as dengo_evolve_<n> does), T[ncells] the cell temperatures of
 * the isothermal synthetic network (NULL otherwise), status[ncells] an
 * optional host array for per-cell status.  dt0 < 0 selects dtf/1e5; z is the
 * redshift (primordial/oxygen).  Returns a ticket >= 0, or a negative code
 * (unknown network, no cells, a patch larger than the arena). */
long long dengo_server_submit(dengo_server *srv, int network, long long ncells, double *state,
        const double *rtol, double *atol, const double *T, int *status, double dtf, double dt0,
        double z);

/* Queues one patch held in device memory of the server's device: d_state,
 * d_rtol, d_atol [ncells][N], d_T[ncells] (isothermal network only) and the
 * optional d_status[ncells] are the caller's, integrated in place with nothing
 * copied (no arena space is used), so a patch's state can stay on the device
 * between its successive requests.  The arrays must stay valid and untouched
 * until the request completes; its completion record carries the status
 * counts when d_status is given.  Each pointer is checked to be device memory that holds the patch
 * (hipPointerGetAttributes, hipMemGetAddressRange).  With cfg->dense_order, a
 * dense-network patch's cells are handed out densest first as for host
 * patches; the order is sorted on the device in arena scratch (4 x 8 bytes per
 * cell plus the sort's own), the only arena space a device patch takes.
 * Returns a ticket >= 0, or a negative code. */
long long dengo_server_submit_device(dengo_server *srv, int network, long long ncells,
        double *d_state, const double *d_rtol, double *d_atol, const double *d_T, int *d_status,
        double dtf, double dt0, double z);

/* Adds a network to a running server: `data` is its data struct (tables set)
 * and `hooks` the dengo_<n>_server_hooks() of the solver library that holds
 * it (any library write_solver emits, loaded in this process).  precision is
 * its sweep's precision (-1: the server's), grid_share its patches' share of
 * the GPU (0: cfg->grid_share).  Call it on the server's device.  Returns the
 * network id to submit with (>= 3), or a negative code. */
int dengo_server_add_network(dengo_server *srv, void *data, const dengo_net_hooks *hooks,
        int precision, int grid_share);

/* Moves up to max completion records into out (completion order); returns
 * how many. */
int dengo_server_poll(dengo_server *srv, dengo_patch_result *out, int max);

/* Waits until every submitted request has completed or timeout_s passed;
 * returns 0, or a negative code on time-out. */
int dengo_server_drain(dengo_server *srv, double timeout_s);

void dengo_server_get_stats(dengo_server *srv, dengo_server_stats *st);

/* Seconds on the server's steady clock (the clock of t_submit/t_done). */
double dengo_server_now(dengo_server *srv);

/* Drains, stops the worker and frees the server's device memory. */
void dengo_server_destroy(dengo_server *srv);

/* Pinned host memory for request arrays (hipHostMalloc); NULL on failure. */
void *dengo_host_alloc(size_t bytes);
void dengo_host_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
