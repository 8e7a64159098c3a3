// Kernels and host implementation of the C ABI, generic over a network.
//
// Kernels (one cell per lane, per-cell state in registers, tables in HBM/L2):
//   k_chem<Net, false, MODE>  one cell per lane, grid covers all cells
//   k_chem<Net, true, MODE>   persistent: resident lanes pull cells from a
//                             device counter (one atomic per wave), one Newton
//                             sweep per loop iteration for every active lane
//   MODE: kFaithful (reference-faithful FP64), kFast64 (block-structured FP64
//         solve), kMixed (FP32 LU + FP64 refinement of the species block)
//   k_eval<Net>               batched RHS / Jacobian / temperature
// A network instantiation file defines the host traits (struct type, table
// fields, field names, blob name) and the extern "C" wrappers.
#pragma once
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <limits>
#include <string>
#include <type_traits>
#include <vector>

#include "be_driver.h"
#include "be_fast.h"
#include "lean_sweep.h"
#include "ma_net.h"
#include "gen_net.h"
#include "dengo_h5.h"
#include "dengo_types.h"

#ifndef MAX_NCELLS
#define MAX_NCELLS 1024   // per-strip scratch length of the emitted data struct
#endif

namespace dengo {

constexpr int kBlock = 64;
enum { kFaithful = 0, kFast64 = 1, kMixed = 2 };

template <class Net, int MODE>
DENGO_D bool cell_start(Cell<Net> &c, long long id, const KParams<Net> &p)
{
    if constexpr (MODE == kFaithful) return start_or_finish(c, id, p);
    else return start_or_finish_fast(c, id, p);
}

template <class Net, int MODE>
DENGO_D bool cell_sweep(Cell<Net> &c, const KParams<Net> &p)
{
    if constexpr (MODE == kFaithful) {
        const int code = sweep_fp64(c, p);
        return code != SW_CONTINUE && end_be(c, code, p);
    } else {
        using LUT = typename std::conditional<MODE == kMixed, float, double>::type;
        const int code = sweep_fast<Net, LUT>(c, p);
        return code != SW_CONTINUE && end_be_fast(c, code, p);
    }
}

template <class Net, bool PERSIST, int MODE>
__global__ void __launch_bounds__(kBlock) k_chem(KParams<Net> p)
{
    Cell<Net> c;
    bool active = false;
    bool exhausted = false;
    if (!PERSIST) {
        const long long id = (long long)blockIdx.x * kBlock + threadIdx.x;
        if (id < p.ncells) active = cell_start<Net, MODE>(c, id, p);
        exhausted = true;
    }
    const int lane = threadIdx.x & 63;
    while (true) {
        if (PERSIST) {
            // lanes without a cell take the next ones, one atomic per wave
            const bool need = !active && !exhausted;
            const unsigned long long mask = __ballot(need);
            if (mask) {
                const int leader = __ffsll((long long)mask) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(p.counter, (unsigned long long)__popcll(mask));
                base = __shfl(base, leader);
                if (need) {
                    const unsigned long long rank =
                        __popcll(mask & ((lane ? (~0ull >> (64 - lane)) : 0ull)));
                    const unsigned long long id = base + rank;
                    if (id < (unsigned long long)p.ncells)
                        active = cell_start<Net, MODE>(c, (long long)id, p);
                    else
                        exhausted = true;
                }
            }
            if (!__any(active || !exhausted)) break;
        } else {
            if (!__any(active)) break;
        }
        if (active && cell_sweep<Net, MODE>(c, p)) {
            finish_cell(c, p);
            active = false;
        }
    }
}

template <class Net>
__global__ void __launch_bounds__(64) k_eval(const double *u_in, long long n,
                                             typename Net::Env env, double *f_out,
                                             double *J_out, double *T_out,
                                             double *rates_out, int *neg_out)
{
    constexpr int N = Net::N, NT = Net::NTAB;
    const long long id = (long long)blockIdx.x * 64 + threadIdx.x;
    if (id >= n) return;
    double u[N], f[N], rs[NT], J[N * N];
    double T;
#pragma unroll
    for (int k = 0; k < N; ++k) u[k] = u_in[id * N + Net::ext(k)];
    const bool ok = Net::rhs(u, f, rs, T, env);
    if (T_out) T_out[id] = T;
    if (rates_out)
#pragma unroll
        for (int t = 0; t < NT; ++t) rates_out[id * NT + t] = rs[t];
    neg_out[id] = ok ? 0 : 1;
    if (!ok) return;
#pragma unroll
    for (int k = 0; k < N; ++k) f_out[id * N + Net::ext(k)] = f[k];
    if (J_out) {
        Net::jacobian(u, rs, T, env, J);
#pragma unroll
        for (int col = 0; col < N; ++col)
#pragma unroll
            for (int row = 0; row < N; ++row)
                J_out[id * N * N + Net::ext(col) * N + Net::ext(row)] = J[col * N + row];
    }
}

// One Newton sweep of BE_chem_solve (BE_chem_solve.C:79-241) for each cell of
// a strip: lane = cell, scaled state u (read only), the cell's updated state
// into u_new, code[cell] = 0 converged, 1 not converged, 2 + FailKind on a
// failed sweep (the host applies the strip-wide consequences), T_out = the
// temperature of the sweep's RHS evaluation.
template <class Net>
__global__ void __launch_bounds__(64) k_strip_sweep(const double *u, double *u_new,
                                                    const double *u0, const double *sc,
                                                    const double *inv, const double *at,
                                                    const double *rt, double dt, long long n,
                                                    typename Net::Env env, double *T_out,
                                                    int *code)
{
    constexpr int N = Net::N;
    const long long id = (long long)blockIdx.x * 64 + threadIdx.x;
    if (id >= n) return;
    Cell<Net> c;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const long long e = id * N + Net::ext(k);
        c.u[k] = u[e];
        c.u0[k] = u0[e];
        c.sc[k] = sc[e];
        c.inv[k] = inv[e];
        c.atol[k] = at[e];
        c.rtol[k] = rt[e];
    }
    c.dt = dt;
    c.sweep = 0;
    c.nsweep = 0;
    c.T = 0.0;
    KParams<Net> p;
    p.env = env;
    int kind = FK_NONE;
    const int r = sweep_fp64(c, p, kind);
    T_out[id] = c.T;
#pragma unroll
    for (int k = 0; k < N; ++k) u_new[id * N + Net::ext(k)] = c.u[k];
    code[id] = r == SW_FAILED ? 2 + kind : (r == SW_CONVERGED ? 0 : 1);
}

// temperature_from_mass_density (rates_and_rate_tables.C.template:635-668):
// species other than ge divided by weight * mh, the network's temperature
// formula, clamped below at 1 K and not above.
template <class Net>
__global__ void __launch_bounds__(64) k_temperature_from_mass(const double *in, long long n,
                                                              double *T_out)
{
    constexpr int N = Net::N;
    const long long id = (long long)blockIdx.x * 64 + threadIdx.x;
    if (id >= n) return;
    double u[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        u[k] = in[id * N + Net::ext(k)];
        if (k != Net::GE) u[k] /= Net::weight(k) * 1.67e-24;
    }
    TableView tb;
    tb.pairs = nullptr;
    tb.nbins = 0;
    tb.dbin = tb.idbin = tb.lb = 0.0;
    tb.T0 = 1.0;
    tb.T1 = __builtin_huge_val();
    T_out[id] = Net::temperature(u, tb);
}

template <int DUMMY>
__global__ void k_reset_counter(unsigned long long *counter)
{
    *counter = 0ull;
}

// Smallest elapsed time over cells (ties: lowest index) and its final dt.
template <int DUMMY>
__global__ void __launch_bounds__(1024) k_min_ttot(const double *ttot, const double *dt,
                                                   long long n, double *out)
{
    __shared__ double sv[1024];
    __shared__ long long si[1024];
    double best = 1e300;
    long long bi = -1;
    for (long long i = threadIdx.x; i < n; i += blockDim.x) {
        const double v = ttot[i];
        if (v < best || (v == best && (bi < 0 || i < bi))) {
            best = v;
            bi = i;
        }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double v = sv[threadIdx.x + s];
            const long long j = si[threadIdx.x + s];
            if (j >= 0 && (si[threadIdx.x] < 0 || v < sv[threadIdx.x] ||
                           (v == sv[threadIdx.x] && j < si[threadIdx.x]))) {
                sv[threadIdx.x] = v;
                si[threadIdx.x] = j;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sv[0];
        out[1] = si[0] >= 0 ? dt[si[0]] : 0.0;
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
int set_error(const char *fmt, ...);   // defined in dengo_common.hip
void set_last_ms(double ms);
const char *last_error_text();

#define DENGO_HIP_OK(expr)                                                               \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return ::dengo::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                      __FILE__, __LINE__);                               \
    } while (0)

struct OrderBufs;

// One integration enqueued by the patch server or captured in a graph: a
// network-independent view of the arguments of <n>_evolve_device with the
// queue counter made explicit.
struct LaunchArgs {
    dengo_options opt;
    double dtf, dt0, z;
    const double *T;              // per-cell temperatures (isothermal dense network)
    double *x;
    const double *rtol;
    double *atol;
    long long n;
    dengo_cell_outputs out;
    hipStream_t stream;
    unsigned long long *counter;  // persistent schedule's work counter
    OrderBufs *ord;               // buffers of opt.order (nullptr: the solver's own)
};

// What the patch server needs of a network: prepare() makes the tables
// resident on the current device and reports the persistent grid that fills
// it for the given options; launch() enqueues one integration.
struct NetHooks {
    int nspecies;
    int isothermal;               // 1: cells carry T[n] instead of an energy unknown
    int (*prepare)(void *data, const dengo_options *opt, int *full_grid);
    // sizes the buffers b (nullptr: the solver's own) for a cell order of n
    // cells and checks the order (synchronises stream)
    int (*prepare_order)(void *data, OrderBufs *b, const long long *order, long long n,
                         void *stream);
    int (*launch)(void *data, const LaunchArgs &a);
    // like prepare, but allocates and uploads nothing: fails unless the
    // tables are already resident on the current device
    int (*resident)(void *data, const dengo_options *opt, int *full_grid);
};
NetHooks hooks_primordial();
NetHooks hooks_oxygen();
NetHooks hooks_synthetic();

// ---------------------------------------------------------------------------
// Cell order (dengo_options.order).  The lane-per-cell networks gather the
// cells into the order's sequence before a launch and scatter them back
// after it, so their kernels see contiguous cells and need no indirection;
// the order must be a permutation of 0..dims-1, which order_check verifies
// (one pass and one synchronisation) before anything is launched.
// ---------------------------------------------------------------------------
struct OrderBufs {
    long long cap = 0;                 // cells
    int n = 0;                         // unknowns per cell of the row buffers
    double *in = nullptr, *rt = nullptr, *at = nullptr;   // [cap][n]
    double *dt = nullptr, *ttot = nullptr, *T = nullptr;  // [cap]
    int *status = nullptr, *nsub = nullptr, *nrej = nullptr, *nsweep = nullptr;
    int *seen = nullptr;
    unsigned long long *bad = nullptr;
};

template <int DUMMY>
__global__ void k_order_check(const long long *order, long long n, int *seen,
                              unsigned long long *bad)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long j = order[i];
    if (j < 0 || j >= n || atomicExch(&seen[j], 1) != 0) atomicAdd(bad, 1ull);
}

// dst row i = src row order[i] (gather) or dst row order[i] = src row i
// (scatter), rows of N values, for up to three arrays at once.
template <int N, bool GATHER>
__global__ void k_order_rows(const long long *order, long long n, const double *a0,
                             const double *a1, const double *a2, double *b0, double *b1,
                             double *b2)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * N) return;
    const long long i = e / N, k = e - i * N;
    const long long o = order[i] * N + k;
    const long long from = GATHER ? o : e, to = GATHER ? e : o;
    if (a0) b0[to] = a0[from];
    if (a1) b1[to] = a1[from];
    if (a2) b2[to] = a2[from];
}

// Per-cell outputs: dst[order[i]] = src[i] for each field dst requests.
template <int DUMMY>
__global__ void k_order_outputs(const long long *order, long long n, dengo_cell_outputs src,
                                dengo_cell_outputs dst)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long j = order[i];
    if (dst.status) dst.status[j] = src.status[i];
    if (dst.nsub) dst.nsub[j] = src.nsub[i];
    if (dst.nrej) dst.nrej[j] = src.nrej[i];
    if (dst.nsweep) dst.nsweep[j] = src.nsweep[i];
    if (dst.dt) dst.dt[j] = src.dt[i];
    if (dst.ttot) dst.ttot[j] = src.ttot[i];
    if (dst.T) dst.T[j] = src.T[i];
}

template <class T>
inline int order_grow(T **p, long long n)
{
    if (*p) DENGO_HIP_OK(hipFree(*p));
    *p = nullptr;
    DENGO_HIP_OK(hipMalloc(p, (size_t)(n > 0 ? n : 1) * sizeof(T)));
    return 0;
}

// Buffers for n cells of nrow unknowns (nrow 0: the permutation check only).
inline int order_alloc(OrderBufs &b, long long n, int nrow)
{
    if (!b.bad) DENGO_HIP_OK(hipMalloc(&b.bad, sizeof(unsigned long long)));
    if (n <= b.cap && nrow <= b.n) return 0;
    const long long cap = n > b.cap ? n : b.cap;
    const int nr = nrow > b.n ? nrow : b.n;
    if (order_grow(&b.seen, cap)) return -1;
    if (nr > 0 &&
        (order_grow(&b.in, cap * nr) || order_grow(&b.rt, cap * nr) || order_grow(&b.at, cap * nr) ||
         order_grow(&b.dt, cap) || order_grow(&b.ttot, cap) || order_grow(&b.T, cap) ||
         order_grow(&b.status, cap) || order_grow(&b.nsub, cap) || order_grow(&b.nrej, cap) ||
         order_grow(&b.nsweep, cap)))
        return -1;
    b.cap = cap;
    b.n = nr;
    return 0;
}

inline void order_free(OrderBufs &b)
{
    void *ptrs[] = {b.in, b.rt, b.at, b.dt, b.ttot, b.T, b.status, b.nsub, b.nrej, b.nsweep,
                    b.seen, b.bad};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    b = OrderBufs();
}

// Fails unless order[0..n) is a permutation of 0..n-1 (synchronises stream).
inline int order_check(OrderBufs &b, const long long *order, long long n, hipStream_t st)
{
    if (n <= 0) return 0;
    DENGO_HIP_OK(hipMemsetAsync(b.seen, 0, (size_t)n * sizeof(int), st));
    DENGO_HIP_OK(hipMemsetAsync(b.bad, 0, sizeof(unsigned long long), st));
    hipLaunchKernelGGL(k_order_check<0>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, order,
                       n, b.seen, b.bad);
    DENGO_HIP_OK(hipGetLastError());
    unsigned long long bad = 0;
    DENGO_HIP_OK(hipMemcpyAsync(&bad, b.bad, sizeof(bad), hipMemcpyDeviceToHost, st));
    DENGO_HIP_OK(hipStreamSynchronize(st));
    if (bad) return set_error("order is not a permutation of 0..%lld (%llu entries repeat or lie "
                              "outside)", n - 1, bad);
    return 0;
}

// The per-cell output fields out requests, in the order buffers.
inline dengo_cell_outputs order_outputs(const OrderBufs &b, const dengo_cell_outputs &out)
{
    dengo_cell_outputs p;
    p.status = out.status ? b.status : nullptr;
    p.nsub = out.nsub ? b.nsub : nullptr;
    p.nrej = out.nrej ? b.nrej : nullptr;
    p.nsweep = out.nsweep ? b.nsweep : nullptr;
    p.dt = out.dt ? b.dt : nullptr;
    p.ttot = out.ttot ? b.ttot : nullptr;
    p.T = out.T ? b.T : nullptr;
    return p;
}

struct DevState {
    int device = -1;
    double2 *pairs = nullptr;
    bool stale = true;
    long long cap = 0;
    double *d_in = nullptr, *d_rtol = nullptr, *d_atol = nullptr;
    double *d_ttot = nullptr, *d_dt = nullptr, *d_T = nullptr;
    int *d_status = nullptr, *d_nsub = nullptr, *d_nrej = nullptr, *d_nsweep = nullptr;
    unsigned long long *d_counter = nullptr;
    double *d_red = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int cu_count = 0;
    // strip solve (be_chem_solve): initial state, scaling and its inverse
    long long strip_cap = 0;
    double *d_u0 = nullptr, *d_sc = nullptr, *d_inv = nullptr, *d_unew = nullptr;
    int *d_code = nullptr;
    OrderBufs ord;                     // cell order: rows and outputs in hand-out order
    // intermediate solutions: -1 follows the network's WRITE_INTERMEDIATE
    int write_im = -1;
};

// Device memory for intermediate-solution records, (N + 3) doubles per cell
// and accepted substep: the records stop at the substep this budget reaches.
constexpr size_t kIntermediateBytes = size_t(1) << 30;

// Intermediate-solution records of one host-array evolve (KParams::traj),
// handed down its own call chain only, never through shared state.
struct Records {
    double *d = nullptr;
    long long cap = 0;
};

inline int check_options(const dengo_options *opt, dengo_options *o)
{
    dengo_default_options(o);
    if (opt) *o = *opt;
    if (o->block != 0 && o->block != kBlock) return set_error("block must be 0 or %d", kBlock);
    if (o->max_substeps <= 0) o->max_substeps = 50000;
    if (o->schedule != DENGO_PER_CELL && o->schedule != DENGO_PERSISTENT)
        return set_error("unknown schedule %d", o->schedule);
    if (o->precision != DENGO_FP64 && o->precision != DENGO_MIXED &&
        o->precision != DENGO_FP64_FAST)
        return set_error("unknown precision %d", o->precision);
    if (o->kernel_form != 0 && o->kernel_form != 1)
        return set_error("kernel_form must be 0 or 1, got %d", o->kernel_form);
    return 0;
}

// Traits (per network): Data (struct), NSPEC, NTABLES, NT, T_LO/T_HI (the
// temperature grid), NZ, Z_LO/Z_HI (the redshift grid), STOP_TIME (<n>_main's
// dtf), WRITE_INTERMEDIATE (dengo_evolve writes <NAME>_intermediate_%06d.h5),
// NAME, table_offsets(), table_lengths() (NT, or NZ for a redshift table),
// table_names() (struct field names), field_names().
template <class Net, class Tr>
struct Host {
    using Data = typename Tr::Data;
    static constexpr int N = Net::N;
    static constexpr int NTAB = Net::NTAB;
    static_assert(Tr::NSPEC == N && Tr::NTABLES == NTAB, "traits do not match the network");

    static DevState *state(Data *d)
    {
        if (!d->device_state) d->device_state = new DevState();
        return (DevState *)d->device_state;
    }

    static const double *table(const Data *d, int t)
    {
        return (const double *)((const char *)d + Tr::table_offsets()[t]);
    }

    static int table_len(int t) { return Tr::table_lengths()[t]; }

    static void set_grid(Data *d, double T0, double T1)
    {
        // rates_and_rate_tables.C.template:25-36
        d->bounds[0] = T0;
        d->bounds[1] = T1;
        d->nbins = Tr::NT - 1;
        d->dbin = (log(d->bounds[1]) - log(d->bounds[0])) / d->nbins;
        d->idbin = (double)(1.0L / d->dbin);
        d->z_bounds[0] = Tr::Z_LO;
        d->z_bounds[1] = Tr::Z_HI;
        d->n_zbins = Tr::NZ - 1;
        if (d->n_zbins > 0) {
            d->d_zbin = (log(d->z_bounds[1] + 1.0) - log(d->z_bounds[0] + 1.0)) / d->n_zbins;
            d->id_zbin = (double)(1.0L / d->d_zbin);
        } else {
            d->d_zbin = 0.0;
            d->id_zbin = 0.0;
        }
    }

    // Value of redshift table r at redshift z (rates_and_rate_tables.C.template:
    // 384-401 and the pi/ph branches of :402-449): linear in ln(1 + z) on the
    // table's bin, the bin clamped to [0, n_zbins - 1]; 0 ("no_photo") unless
    // z_bounds[0] <= z < z_bounds[1].
    static double z_rate(const Data *d, const double *r, double z)
    {
        if (!(z >= d->z_bounds[0] && z < d->z_bounds[1])) return 0.0;
        const double lbz = log(d->z_bounds[0] + 1.0);
        int zbin_id = (int)(d->id_zbin * (log(z + 1.0) - lbz));
        if (zbin_id <= 0) zbin_id = 0;
        else if (zbin_id >= d->n_zbins) zbin_id = d->n_zbins - 1;
        const double z1 = (lbz + (zbin_id) * d->d_zbin);
        const double z2 = (lbz + (zbin_id + 1) * d->d_zbin);
        const double zdef = (log(z + 1.0) - z1) / (z2 - z1);
        return r[zbin_id] + zdef * (r[zbin_id + 1] - r[zbin_id]);
    }

    static int ensure_device(Data *d, DevState *s)
    {
        int dev = 0;
        DENGO_HIP_OK(hipGetDevice(&dev));
        if (s->device != dev) {
            if (s->device >= 0)
                return set_error("data struct was set up on device %d, current is %d", s->device,
                                 dev);
            s->device = dev;
            DENGO_HIP_OK(hipMalloc(&s->d_counter, sizeof(unsigned long long)));
            DENGO_HIP_OK(hipMalloc(&s->d_red, 2 * sizeof(double)));
            DENGO_HIP_OK(hipEventCreate(&s->ev0));
            DENGO_HIP_OK(hipEventCreate(&s->ev1));
            hipDeviceProp_t prop;
            DENGO_HIP_OK(hipGetDeviceProperties(&prop, dev));
            s->cu_count = prop.multiProcessorCount;
        }
        if (s->stale || !s->pairs) {
            if (d->nbins != Tr::NT - 1) return set_error("tables not set (nbins=%d)", d->nbins);
            const int nb = d->nbins;
            std::vector<double2> host((size_t)nb * NTAB);
            for (int t = 0; t < NTAB; ++t) {
                const double *r = table(d, t);
                for (int b = 0; b < nb; ++b)
                    host[(size_t)b * NTAB + t] = Net::ztab(t) >= 0 ? make_double2(0.0, 0.0)
                                                                   : make_double2(r[b], r[b + 1]);
            }
            if (!s->pairs) DENGO_HIP_OK(hipMalloc(&s->pairs, host.size() * sizeof(double2)));
            DENGO_HIP_OK(hipMemcpy(s->pairs, host.data(), host.size() * sizeof(double2),
                                   hipMemcpyHostToDevice));
            s->stale = false;
        }
        return 0;
    }

    template <class T>
    static int grow(T **p, long long n)
    {
        if (*p) DENGO_HIP_OK(hipFree(*p));
        *p = nullptr;
        DENGO_HIP_OK(hipMalloc(p, (size_t)n * sizeof(T)));
        return 0;
    }

    static int ensure_scratch(DevState *s, long long n)
    {
        if (n <= s->cap) return 0;
        if (grow(&s->d_in, n * N) || grow(&s->d_rtol, n * N) || grow(&s->d_atol, n * N) ||
            grow(&s->d_ttot, n) || grow(&s->d_dt, n) || grow(&s->d_T, n) ||
            grow(&s->d_status, n) || grow(&s->d_nsub, n) || grow(&s->d_nrej, n) ||
            grow(&s->d_nsweep, n))
            return -1;
        s->cap = n;
        return 0;
    }

    static typename Net::Env env(const Data *d, const DevState *s, double z)
    {
        typename Net::Env e;
        e.tb.pairs = s->pairs;
        e.tb.nbins = d->nbins;
        e.tb.dbin = d->dbin;
        e.tb.idbin = d->idbin;
        e.tb.lb = log(d->bounds[0]);
        e.tb.T0 = d->bounds[0];
        e.tb.T1 = d->bounds[1];
        e.zp1 = z + 1.0;
        e.comp4 = pow(e.zp1, 4);
        e.z = z;
        for (int k = 0; k < prim::kMaxZTables; ++k) e.zr[k] = 0.0;
        for (int t = 0; t < NTAB; ++t)
            if (Net::ztab(t) >= 0) e.zr[Net::ztab(t)] = z_rate(d, table(d, t), z);
        return e;
    }

    static int read_blob(const char *path, Data *d)
    {
        FILE *f = fopen(path, "rb");
        if (!f) return set_error("cannot open table file %s", path);
        char magic[8];
        int32_t nt = 0, nT = 0;
        double hdr[4];
        char pad[64 - 8 - 8 - 32];
        int ok = fread(magic, 1, 8, f) == 8 && fread(&nt, 4, 1, f) == 1 &&
                 fread(&nT, 4, 1, f) == 1 && fread(hdr, 8, 4, f) == 4 &&
                 fread(pad, 1, sizeof(pad), f) == sizeof(pad);
        if (!ok || memcmp(magic, "DENGOTB1", 8) != 0) {
            fclose(f);
            return set_error("%s is not a DENGOTB1 table blob", path);
        }
        if (nt != NTAB || nT != Tr::NT) {
            fclose(f);
            return set_error("%s holds %d tables x %d bins, the struct needs %d x %d", path, nt,
                             nT, NTAB, Tr::NT);
        }
        std::vector<double> row((size_t)nT);
        for (int t = 0; t < nt; ++t) {
            // every row holds nT values; a redshift table keeps its first NZ
            if (fread(row.data(), 8, nT, f) != (size_t)nT) {
                fclose(f);
                return set_error("%s: short read", path);
            }
            memcpy((double *)table(d, t), row.data(), sizeof(double) * table_len(t));
        }
        fclose(f);
        set_grid(d, hdr[0], hdr[1]);
        state(d)->stale = true;
        return 0;
    }

    // The reference's <n>_tables.h5 (dengo/chemical_network.py:310-318; read
    // by rates_and_rate_tables.C.template:278-304): one FP64 dataset of NT
    // values per table, named after the struct field without its r_/c_
    // prefix; the temperature grid is the network's (:25-29).
    static int read_h5(const char *path, Data *d)
    {
        const dengo_h5_file f = dengo_h5_open(path, 0);
        if (f < 0) return -1;
        for (int t = 0; t < NTAB; ++t) {
            char name[128];
            snprintf(name, sizeof(name), "/%s", Tr::table_names()[t] + 2);
            if (dengo_h5_read(f, name, (double *)table(d, t), table_len(t)) != 0) {
                char why[256];
                snprintf(why, sizeof(why), "%s", last_error_text());
                dengo_h5_close(f);
                return set_error("%s: %s", path, why);
            }
        }
        dengo_h5_close(f);
        set_grid(d, Tr::T_LO, Tr::T_HI);
        state(d)->stale = true;
        return 0;
    }

    // A table file of either kind, told apart by its first bytes.
    static int read_tables(const char *path, Data *d)
    {
        return dengo_h5_is_hdf5(path) ? read_h5(path, d) : read_blob(path, d);
    }

    // <n>_<suffix>.h5 when it is an HDF5 file in the working directory (the
    // reference's file) and libhdf5 can be loaded, else <n>_<suffix>.bin.
    static const char *default_file(const char *n, const char *suffix, char *buf, size_t len)
    {
        snprintf(buf, len, "%s_%s.h5", n, suffix);
        if (dengo_h5_is_hdf5(buf) && dengo_hdf5_available()) return buf;
        snprintf(buf, len, "%s_%s.bin", n, suffix);
        return buf;
    }

    // Field files of this backend without HDF5 ("DENGOIC1"): a 40-byte header
    // (magic, int32 nfields, int32 0, int64 ncells, FP64 time, FP64 timestep),
    // nfields names of 32 bytes, then FP64 [nfields][ncells].
    static int read_fields_bin(const char *path, std::vector<std::string> &names,
                               std::vector<double> &vals, long long &ncells, double *hdr2)
    {
        FILE *f = fopen(path, "rb");
        if (!f) return set_error("cannot open %s", path);
        char magic[8];
        int32_t nf = 0, pad = 0;
        long long nc = 0;
        double tt[2];
        const bool ok = fread(magic, 1, 8, f) == 8 && fread(&nf, 4, 1, f) == 1 &&
                        fread(&pad, 4, 1, f) == 1 && fread(&nc, 8, 1, f) == 1 &&
                        fread(tt, 8, 2, f) == 2 && memcmp(magic, "DENGOIC1", 8) == 0 &&
                        nf > 0 && nf < 4096 && nc >= 0;
        if (!ok) {
            fclose(f);
            return set_error("%s is not a DENGOIC1 field file", path);
        }
        names.assign(nf, std::string());
        for (int i = 0; i < nf; ++i) {
            char nm[33] = {0};
            if (fread(nm, 1, 32, f) != 32) {
                fclose(f);
                return set_error("%s: short read", path);
            }
            names[i] = nm;
        }
        vals.resize((size_t)nf * nc);
        if (nc > 0 && fread(vals.data(), 8, vals.size(), f) != vals.size()) {
            fclose(f);
            return set_error("%s: short read", path);
        }
        fclose(f);
        ncells = nc;
        if (hdr2) {
            hdr2[0] = tt[0];
            hdr2[1] = tt[1];
        }
        return 0;
    }

    static int write_fields_bin(const char *path, const std::vector<std::string> &names,
                                const std::vector<const double *> &cols, long long ncells,
                                double time, double timestep)
    {
        FILE *f = fopen(path, "wb");
        if (!f) return set_error("cannot create %s", path);
        const int32_t nf = (int32_t)names.size(), pad = 0;
        const double tt[2] = {time, timestep};
        bool ok = fwrite("DENGOIC1", 1, 8, f) == 8 && fwrite(&nf, 4, 1, f) == 1 &&
                  fwrite(&pad, 4, 1, f) == 1 && fwrite(&ncells, 8, 1, f) == 1 &&
                  fwrite(tt, 8, 2, f) == 2;
        for (int i = 0; ok && i < nf; ++i) {
            char nm[32] = {0};
            strncpy(nm, names[i].c_str(), 31);
            ok = fwrite(nm, 1, 32, f) == 32;
        }
        for (int i = 0; ok && i < nf; ++i)
            ok = fwrite(cols[i], 8, (size_t)ncells, f) == (size_t)ncells;
        fclose(f);
        return ok ? 0 : set_error("%s: short write", path);
    }

    // Initial conditions AoS [dims][N] from <n>_initial_conditions.h5 (one
    // dataset per species; the cell count from the energy field, as
    // rates_and_rate_tables.C.template:63-107) or the DENGOIC1 file.
    static int read_ics(const char *n, std::vector<double> &ics, long long &dims)
    {
        char buf[512];
        const char *path = default_file(n, "initial_conditions", buf, sizeof(buf));
        const char *const *fields = Tr::field_names();
        if (dengo_h5_is_hdf5(path)) {
            const dengo_h5_file f = dengo_h5_open(path, 0);
            if (f < 0) return -1;
            char name[128];
            snprintf(name, sizeof(name), "/%s", fields[Net::GE]);
            dims = dengo_h5_length(f, name);
            if (dims < 0) {
                dengo_h5_close(f);
                return set_error("%s has no energy field %s", path, name);
            }
            ics.assign((size_t)dims * N, 0.0);
            std::vector<double> col((size_t)dims);
            for (int i = 0; i < N; ++i) {
                snprintf(name, sizeof(name), "/%s", fields[i]);
                if (dengo_h5_read(f, name, col.data(), dims) != 0) {
                    char why[256];
                    snprintf(why, sizeof(why), "%s", last_error_text());
                    dengo_h5_close(f);
                    return set_error("%s: %s", path, why);
                }
                for (long long j = 0; j < dims; ++j) ics[(size_t)j * N + i] = col[j];
            }
            dengo_h5_close(f);
            return 0;
        }
        std::vector<std::string> names;
        std::vector<double> vals;
        if (read_fields_bin(path, names, vals, dims, nullptr) != 0) return -1;
        ics.assign((size_t)dims * N, 0.0);
        for (int i = 0; i < N; ++i) {
            int k = -1;
            for (size_t q = 0; q < names.size(); ++q)
                if (names[q] == fields[i]) k = (int)q;
            if (k < 0) return set_error("%s has no field %s", path, fields[i]);
            for (long long j = 0; j < dims; ++j) ics[(size_t)j * N + i] = vals[(size_t)k * dims + j];
        }
        return 0;
    }

    // <n>_main (rates_and_rate_tables.C.template:56-147): tables and initial
    // conditions from the working directory, atol = 1e-9 x IC, rtol = 1e-9,
    // dtf = the network's stop time, dt = z = -1, dengo_evolve, then
    // <n>_solution.h5 (.bin without HDF5): one dataset per species, /T, and
    // the root attributes time (the returned ttot) and timestep (the final dt).
    static int main_driver(const char *n)
    {
        char buf[512];
        Data *d = setup_from(default_file(n, "tables", buf, sizeof(buf)), NULL, NULL);
        if (!d) return 1;
        std::vector<double> input;
        long long dims = 0;
        if (read_ics(n, input, dims) != 0) {
            fprintf(stderr, "%s_main: %s\n", n, last_error_text());
            free_data(d);
            return 1;
        }
        std::vector<double> rtol(input.size(), 1e-9), atol(input.size()), T((size_t)dims);
        for (size_t i = 0; i < input.size(); ++i) atol[i] = input[i] * 1e-9;
        dengo_options o;
        dengo_default_options(&o);
        double dt = -1.0;
        const double ttot = evolve_host(d, &o, Tr::STOP_TIME, &dt, -1.0, input.data(), rtol.data(),
                                        atol.data(), dims, NULL, NULL, NULL, NULL, NULL, T.data());
        free_data(d);
        if (ttot < 0) {
            fprintf(stderr, "%s_main: %s\n", n, last_error_text());
            return 1;
        }
        std::vector<std::vector<double>> cols(N, std::vector<double>((size_t)dims));
        for (int i = 0; i < N; ++i)
            for (long long j = 0; j < dims; ++j) cols[i][j] = input[(size_t)j * N + i];
        const char *const *fields = Tr::field_names();
        int rc = 0;
        if (dengo_hdf5_available()) {
            snprintf(buf, sizeof(buf), "%s_solution.h5", n);
            const dengo_h5_file f = dengo_h5_open(buf, 1);
            if (f < 0) rc = -1;
            char name[128];
            for (int i = 0; rc == 0 && i < N; ++i) {
                snprintf(name, sizeof(name), "/%s", fields[i]);
                rc = dengo_h5_write(f, name, cols[i].data(), dims);
            }
            if (rc == 0) rc = dengo_h5_write(f, "/T", T.data(), dims);
            if (rc == 0) rc = dengo_h5_set_attribute(f, "time", ttot);
            if (rc == 0) rc = dengo_h5_set_attribute(f, "timestep", dt);
            if (f >= 0 && dengo_h5_close(f) != 0) rc = -1;
        } else {
            snprintf(buf, sizeof(buf), "%s_solution.bin", n);
            std::vector<std::string> names(fields, fields + N);
            names.push_back("T");
            std::vector<const double *> p;
            for (int i = 0; i < N; ++i) p.push_back(cols[i].data());
            p.push_back(T.data());
            rc = write_fields_bin(buf, names, p, dims, ttot, dt);
        }
        if (rc != 0) {
            fprintf(stderr, "%s_main: %s\n", n, last_error_text());
            return 1;
        }
        return 0;
    }

    static void field_names(int *NumberOfFields, char ***FieldNames)
    {
        if (FieldNames != NULL && NumberOfFields != NULL) {
            NumberOfFields[0] = N;
            FieldNames[0] = new char *[N];
            for (int i = 0; i < N; ++i) FieldNames[0][i] = strdup(Tr::field_names()[i]);
        }
    }

    static Data *alloc() { return (Data *)calloc(1, sizeof(Data)); }

    static int set_tables(Data *d, const double *tables, int ntables, int nT, double T0, double T1)
    {
        if (!d || !tables) return set_error("null argument");
        if (ntables != NTAB || nT != Tr::NT)
            return set_error("expected %d tables x %d bins, got %d x %d", NTAB, Tr::NT, ntables, nT);
        if (!(T0 > 0) || !(T1 > T0)) return set_error("bad temperature bounds");
        // rows of nT values; a redshift table keeps the first NZ of its row
        for (int t = 0; t < ntables; ++t)
            memcpy((double *)table(d, t), tables + (size_t)t * nT, sizeof(double) * table_len(t));
        set_grid(d, T0, T1);
        state(d)->stale = true;
        return 0;
    }

    static Data *setup_from(const char *path, int *NumberOfFields, char ***FieldNames)
    {
        Data *d = alloc();
        if (!d) return NULL;
        if (read_tables(path, d) != 0) {
            fprintf(stderr, "%s\n", last_error_text());
            free(d);
            return NULL;
        }
        field_names(NumberOfFields, FieldNames);
        return d;
    }

    static void free_data(Data *d)
    {
        if (!d) return;
        DevState *s = (DevState *)d->device_state;
        if (s) {
            void *ptrs[] = {s->pairs, s->d_in, s->d_rtol, s->d_atol, s->d_ttot, s->d_dt, s->d_T,
                            s->d_status, s->d_nsub, s->d_nrej, s->d_nsweep, s->d_counter,
                            s->d_red, s->d_u0, s->d_sc, s->d_inv, s->d_unew, s->d_code};
            for (void *q : ptrs)
                if (q) (void)hipFree(q);
            if (s->ev0) (void)hipEventDestroy(s->ev0);
            if (s->ev1) (void)hipEventDestroy(s->ev1);
            order_free(s->ord);
            delete s;
        }
        free(d);
    }

    // The FP64-fast sweep runs as the register-lean kernel (lean_sweep.h, two
    // waves per SIMD) for every network with a lean adapter (lean::LeanOf)
    // whose lean kernel compiled without scratch and fits at least six 64-lane
    // workgroups per CU; otherwise, and for the other precisions, as the
    // register-resident k_chem.  The two forms of the primordial network are
    // bit-identical.  dengo_options.kernel_form = 1, or building with
    // -DDENGO_LEAN=0, selects k_chem.
#ifndef DENGO_LEAN
#define DENGO_LEAN 1
#endif
    template <class LT>
    static bool lean_fits()
    {
        // per process (the kernels are the same on every device); threads that
        // race here compute the same answer
        static std::atomic<int> fits{-1};
        if (fits.load() < 0) {
            int ok = 1;
            for (const void *k : {(const void *)k_chem_lean<LT, true>,
                                  (const void *)k_chem_lean<LT, false>}) {
                hipFuncAttributes a;
                int per_cu = 0;
                if (hipFuncGetAttributes(&a, k) != hipSuccess || a.localSizeBytes != 0 ||
                    hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kBlock, 0) !=
                        hipSuccess || per_cu < 6)
                    ok = 0;
            }
            (void)hipGetLastError();
            fits.store(ok);
        }
        return fits.load() == 1;
    }

    template <int MODE>
    static void kernels(void (**kp)(KParams<Net>), void (**kc)(KParams<Net>), int form)
    {
        using LT = typename lean::LeanOf<Net>::type;
        if constexpr (DENGO_LEAN && MODE == kFast64 && !std::is_void<LT>::value) {
            if (form == 0 && lean_fits<LT>()) {
                *kp = k_chem_lean<LT, true>;
                *kc = k_chem_lean<LT, false>;
                return;
            }
        }
        *kp = k_chem<Net, true, MODE>;
        *kc = k_chem<Net, false, MODE>;
    }

    // Workgroups of a persistent launch that fill the GPU at full occupancy.
    static int full_grid(DevState *s, const dengo_options &o, int *grid)
    {
        void (*kp)(KParams<Net>) = nullptr;
        void (*kc)(KParams<Net>) = nullptr;
        if (o.precision == DENGO_FP64) kernels<kFaithful>(&kp, &kc, o.kernel_form);
        else if (o.precision == DENGO_FP64_FAST) kernels<kFast64>(&kp, &kc, o.kernel_form);
        else kernels<kMixed>(&kp, &kc, o.kernel_form);
        int per_cu = 0;
        DENGO_HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kp, kBlock, 0));
        if (per_cu < 1) per_cu = 1;
        *grid = per_cu * s->cu_count;
        return 0;
    }

    // Enqueues the integration on stream.  counter: the persistent schedule's
    // work counter (one per concurrently running launch); events: bracket the
    // kernel with the struct's timing events (read by evolve_device).
    static int launch(Data *d, DevState *s, const dengo_options &o, double dtf, double dt0,
                      double z, double *d_input, const double *d_rtol, double *d_atol,
                      long long dims, const dengo_cell_outputs &out, hipStream_t stream,
                      unsigned long long *counter, bool events, OrderBufs *ob = nullptr,
                      const Records *rec = nullptr)
    {
        if (!o.order || dims <= 0)
            return launch_cells(d, s, o, dtf, dt0, z, d_input, d_rtol, d_atol, dims, out, stream,
                                counter, events, rec);
        // cells in hand-out order: gather, integrate, scatter back (the
        // buffers were allocated and the order checked by the caller)
        OrderBufs &b = ob ? *ob : s->ord;
        if (b.cap < dims || b.n < N) return set_error("cell order buffers not prepared");
        dengo_options oc = o;
        oc.order = nullptr;
        const dengo_cell_outputs pout = order_outputs(b, out);
        const unsigned gr = (unsigned)((dims * N + 255) / 256), gc = (unsigned)((dims + 255) / 256);
        // the timed interval covers the gather and the scatters too
        if (events) DENGO_HIP_OK(hipEventRecord(s->ev0, stream));
        hipLaunchKernelGGL((k_order_rows<N, true>), dim3(gr), dim3(256), 0, stream, o.order, dims,
                           d_input, d_rtol, d_atol, b.in, b.rt, b.at);
        DENGO_HIP_OK(hipGetLastError());
        if (launch_cells(d, s, oc, dtf, dt0, z, b.in, b.rt, b.at, dims, pout, stream, counter,
                         false))
            return -1;
        hipLaunchKernelGGL((k_order_rows<N, false>), dim3(gr), dim3(256), 0, stream, o.order, dims,
                           b.in, b.at, nullptr, d_input, d_atol, nullptr);
        DENGO_HIP_OK(hipGetLastError());
        hipLaunchKernelGGL(k_order_outputs<0>, dim3(gc), dim3(256), 0, stream, o.order, dims, pout,
                           out);
        DENGO_HIP_OK(hipGetLastError());
        if (events) DENGO_HIP_OK(hipEventRecord(s->ev1, stream));
        return 0;
    }

    static int launch_cells(Data *d, DevState *s, const dengo_options &o, double dtf, double dt0,
                            double z, double *d_input, const double *d_rtol, double *d_atol,
                            long long dims, const dengo_cell_outputs &out, hipStream_t stream,
                            unsigned long long *counter, bool events,
                            const Records *rec = nullptr)
    {
        KParams<Net> p;
        p.input = d_input;
        p.rtol = d_rtol;
        p.atol = d_atol;
        p.ncells = dims;
        p.env = env(d, s, z);
        p.dtf = dtf;
        p.dt_in = dt0;
        p.input_is_number = o.input_is_number;
        p.max_substeps = o.max_substeps;
        p.status = out.status;
        p.nsub = out.nsub;
        p.nrej = out.nrej;
        p.nsweep = out.nsweep;
        p.dt_out = out.dt;
        p.ttot_out = out.ttot;
        p.T_out = out.T;
        p.counter = counter;
        // intermediate records come from the reference-faithful sweep only
        p.traj = rec && o.precision == DENGO_FP64 ? rec->d : nullptr;
        p.traj_cap = rec ? rec->cap : 0;
        if (dims <= 0) return 0;
        void (*kp)(KParams<Net>) = nullptr;
        void (*kc)(KParams<Net>) = nullptr;
        if (o.precision == DENGO_FP64) kernels<kFaithful>(&kp, &kc, o.kernel_form);
        else if (o.precision == DENGO_FP64_FAST) kernels<kFast64>(&kp, &kc, o.kernel_form);
        else kernels<kMixed>(&kp, &kc, o.kernel_form);
        if (events) DENGO_HIP_OK(hipEventRecord(s->ev0, stream));
        if (o.schedule == DENGO_PERSISTENT) {
            int grid = o.grid;
            if (grid <= 0 && full_grid(s, o, &grid)) return -1;
            const long long need = (dims + kBlock - 1) / kBlock;
            if ((long long)grid > need) grid = (int)need;
            hipLaunchKernelGGL(k_reset_counter<0>, dim3(1), dim3(1), 0, stream, counter);
            DENGO_HIP_OK(hipGetLastError());
            hipLaunchKernelGGL(kp, dim3(grid), dim3(kBlock), 0, stream, p);
        } else {
            const long long blocks = (dims + kBlock - 1) / kBlock;
            if (blocks > 0x7fffffffLL) return set_error("too many cells: %lld", dims);
            hipLaunchKernelGGL(kc, dim3((unsigned)blocks), dim3(kBlock), 0, stream, p);
        }
        DENGO_HIP_OK(hipGetLastError());
        if (events) DENGO_HIP_OK(hipEventRecord(s->ev1, stream));
        return 0;
    }

    static int evolve_device(Data *d, const dengo_options *opt, double dtf, double dt0, double z,
                             double *d_input, const double *d_rtol, double *d_atol,
                             long long dims, const dengo_cell_outputs *outputs,
                             double *ttot_min, double *dt_min, void *stream,
                             const Records *rec = nullptr)
    {
        if (!d || !d_input || !d_rtol || !d_atol) return set_error("null argument");
        if (dims < 0) return set_error("negative cell count");
        dengo_options o;
        if (check_options(opt, &o)) return -1;
        DevState *s = state(d);
        if (ensure_device(d, s)) return -1;
        dengo_cell_outputs out;
        memset(&out, 0, sizeof(out));
        if (outputs) out = *outputs;
        hipStream_t st = (hipStream_t)stream;
        if (ttot_min && (!out.ttot || !out.dt)) {
            if (ensure_scratch(s, dims)) return -1;
            if (!out.ttot) out.ttot = s->d_ttot;
            if (!out.dt) out.dt = s->d_dt;
        }
        if (o.order && dims > 0 &&
            (order_alloc(s->ord, dims, N) || order_check(s->ord, o.order, dims, st)))
            return -1;
        d->current_z = z;
        if (launch(d, s, o, dtf, dt0, z, d_input, d_rtol, d_atol, dims, out, st, s->d_counter, true,
                   nullptr, rec))
            return -1;
        if (ttot_min) {
            if (dims > 0) {
                hipLaunchKernelGGL(k_min_ttot<0>, dim3(1), dim3(1024), 0, st, out.ttot, out.dt,
                                   dims, s->d_red);
                DENGO_HIP_OK(hipGetLastError());
                double red[2];
                DENGO_HIP_OK(hipMemcpyAsync(red, s->d_red, sizeof(red), hipMemcpyDeviceToHost, st));
                DENGO_HIP_OK(hipStreamSynchronize(st));
                *ttot_min = red[0];
                if (dt_min) *dt_min = red[1];
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, s->ev0, s->ev1) == hipSuccess) set_last_ms(ms);
            } else {
                *ttot_min = 0.0;
                if (dt_min) *dt_min = dt0;
            }
        }
        return 0;
    }

    static double evolve_host(Data *d, const dengo_options *opt, double dtf, double *dt, double z,
                              double *input, double *rtol, double *atol, long long dims,
                              int *status, int *nsub, int *nrej, int *nsweep, double *dt_cell,
                              double *T_cell)
    {
        if (!d || !dt || !input || !rtol || !atol) {
            set_error("null argument");
            return -1.0;
        }
        if (dims <= 0) return 0.0;
        DevState *s = state(d);
        if (ensure_device(d, s) || ensure_scratch(s, dims)) return -1.0;
        if (writes_intermediate(s)) {
            Records rec;
            const double t = evolve_host_intermediate(d, s, opt, dtf, dt, z, input, rtol, atol,
                                                      dims, status, nsub, nrej, nsweep, dt_cell,
                                                      T_cell, rec);
            if (rec.d) (void)hipFree(rec.d);
            return t;
        }
        return evolve_host_plain(d, s, opt, dtf, dt, z, input, rtol, atol, dims, status, nsub,
                                 nrej, nsweep, dt_cell, T_cell);
    }

    static bool writes_intermediate(const DevState *s)
    {
        return s->write_im < 0 ? Tr::WRITE_INTERMEDIATE : s->write_im != 0;
    }

    static int set_write_intermediate(Data *d, int on)
    {
        if (!d) return set_error("null argument");
        state(d)->write_im = on ? 1 : 0;
        return 0;
    }

    // dengo_evolve with <n>_intermediate_%06d.h5 (rates_and_rate_tables.C.template:
    // 212-241): the faithful sweep records, per cell and accepted substep, the
    // state before it, the last sweep's T, the elapsed time and the next dt;
    // file k then holds every cell's k-th record.  The reference's cells share
    // their substeps; here each cell has its own, so a cell with fewer than k
    // repeats its last record (NaN without one), the root attributes time and
    // timestep are those of the cell with the least elapsed time (the rule of
    // <n>_solution.h5), and cell_time / cell_timestep hold every cell's.
    static double evolve_host_intermediate(Data *d, DevState *s, const dengo_options *opt,
                                           double dtf, double *dt, double z, double *input,
                                           double *rtol, double *atol, long long dims,
                                           int *status, int *nsub, int *nrej, int *nsweep,
                                           double *dt_cell, double *T_cell, Records &rc)
    {
        dengo_options o;
        if (check_options(opt, &o)) return -1.0;
        o.precision = DENGO_FP64;
        o.order = nullptr;
        const size_t rec = (size_t)dims * (N + 3) * sizeof(double);
        long long cap = (long long)(kIntermediateBytes / rec);
        if (cap > o.max_substeps) cap = o.max_substeps;
        if (cap < 1) {
            set_error("%lld cells do not fit one intermediate record in %zu bytes", dims,
                      kIntermediateBytes);
            return -1.0;
        }
        if (hipMalloc(&rc.d, (size_t)cap * rec) != hipSuccess) {
            rc.d = nullptr;
            set_error("cannot allocate %zu bytes of intermediate records", (size_t)cap * rec);
            return -1.0;
        }
        rc.cap = cap;
        std::vector<int> ns((size_t)dims);
        const double ttot = evolve_host_plain(d, s, &o, dtf, dt, z, input, rtol, atol, dims, status,
                                              ns.data(), nrej, nsweep, dt_cell, T_cell, &rc);
        if (ttot < 0) return ttot;
        if (nsub) memcpy(nsub, ns.data(), (size_t)dims * sizeof(int));
        long long K = 0;
        bool cut = false;
        for (long long j = 0; j < dims; ++j) {
            const long long m = ns[j] < cap ? ns[j] : cap;
            cut = cut || ns[j] > cap;
            if (m > K) K = m;
        }
        std::vector<double> recs((size_t)K * dims * (N + 3));
        if (K > 0 &&
            hipMemcpy(recs.data(), rc.d, recs.size() * sizeof(double),
                      hipMemcpyDeviceToHost) != hipSuccess) {
            set_error("device-to-host copy failed");
            return -1.0;
        }
        if (cut)
            fprintf(stderr, "dengo_evolve_%s: intermediate files stop at substep %lld (%zu bytes "
                            "of records)\n", Tr::NAME, cap, kIntermediateBytes);
        if (write_intermediate_files(recs.data(), ns.data(), cap, K, dims) != 0) return -1.0;
        return ttot;
    }

    static int write_intermediate_files(const double *recs, const int *ns, long long cap,
                                        long long K, long long dims)
    {
        const double nan = std::numeric_limits<double>::quiet_NaN();
        std::vector<std::vector<double>> cols(N + 3, std::vector<double>((size_t)dims));
        const char *const *fields = Tr::field_names();
        std::vector<std::string> names(fields, fields + N);
        names.push_back("T");
        char path[512], name[128];
        for (long long k = 1; k <= K; ++k) {
            double tmin = nan, dtmin = nan;
            for (long long j = 0; j < dims; ++j) {
                const long long m = ns[j] < cap ? ns[j] : cap;
                if (m == 0) {
                    for (int i = 0; i < N + 3; ++i) cols[i][j] = nan;
                    continue;
                }
                const double *r = recs + ((size_t)((k <= m ? k : m) - 1) * dims + j) * (N + 3);
                for (int i = 0; i < N + 3; ++i) cols[i][j] = r[i];
                if (!(r[N + 1] >= tmin)) {
                    tmin = r[N + 1];
                    dtmin = r[N + 2];
                }
            }
            int rc = 0;
            if (dengo_hdf5_available()) {
                snprintf(path, sizeof(path), "%s_intermediate_%06lld.h5", Tr::NAME, k);
                const dengo_h5_file f = dengo_h5_open(path, 1);
                if (f < 0) return -1;
                for (int i = 0; rc == 0 && i <= N; ++i) {
                    snprintf(name, sizeof(name), "/%s", names[i].c_str());
                    rc = dengo_h5_write(f, name, cols[i].data(), dims);
                }
                if (rc == 0) rc = dengo_h5_set_attribute(f, "time", tmin);
                if (rc == 0) rc = dengo_h5_set_attribute(f, "timestep", dtmin);
                if (rc == 0) rc = dengo_h5_set_attribute_array(f, "cell_time", cols[N + 1].data(), dims);
                if (rc == 0)
                    rc = dengo_h5_set_attribute_array(f, "cell_timestep", cols[N + 2].data(), dims);
                if (dengo_h5_close(f) != 0) rc = -1;
            } else {
                snprintf(path, sizeof(path), "%s_intermediate_%06lld.bin", Tr::NAME, k);
                std::vector<const double *> p;
                for (int i = 0; i <= N; ++i) p.push_back(cols[i].data());
                rc = write_fields_bin(path, names, p, dims, tmin, dtmin);
            }
            if (rc != 0) return -1;
        }
        return 0;
    }

    static double evolve_host_plain(Data *d, DevState *s, const dengo_options *opt, double dtf,
                                    double *dt, double z, double *input, double *rtol,
                                    double *atol, long long dims, int *status, int *nsub,
                                    int *nrej, int *nsweep, double *dt_cell, double *T_cell,
                                    const Records *rec = nullptr)
    {
        const size_t bytes = (size_t)dims * N * sizeof(double);
        if (hipMemcpy(s->d_in, input, bytes, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s->d_rtol, rtol, bytes, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s->d_atol, atol, bytes, hipMemcpyHostToDevice) != hipSuccess) {
            set_error("host-to-device copy failed");
            return -1.0;
        }
        dengo_cell_outputs out;
        out.status = s->d_status;
        out.nsub = s->d_nsub;
        out.nrej = s->d_nrej;
        out.nsweep = s->d_nsweep;
        out.dt = s->d_dt;
        out.ttot = s->d_ttot;
        out.T = s->d_T;
        double tmin = 0.0, dtm = 0.0;
        if (evolve_device(d, opt, dtf, *dt, z, s->d_in, s->d_rtol, s->d_atol, dims, &out, &tmin,
                          &dtm, NULL, rec) != 0)
            return -1.0;
        struct Copy { void *dst; const void *src; size_t n; };
        const long long nts = dims < MAX_NCELLS ? dims : MAX_NCELLS;
        const Copy cps[] = {{input, s->d_in, bytes}, {atol, s->d_atol, bytes},
                            {status, s->d_status, dims * sizeof(int)},
                            {nsub, s->d_nsub, dims * sizeof(int)},
                            {nrej, s->d_nrej, dims * sizeof(int)},
                            {nsweep, s->d_nsweep, dims * sizeof(int)},
                            {dt_cell, s->d_dt, dims * sizeof(double)},
                            {T_cell, s->d_T, dims * sizeof(double)},
                            // the reference leaves the last strip's temperatures in data->Ts
                            {d->Ts, s->d_T, nts * sizeof(double)}};
        for (const Copy &c : cps) {
            if (c.dst && hipMemcpy(c.dst, c.src, c.n, hipMemcpyDeviceToHost) != hipSuccess) {
                set_error("device-to-host copy failed");
                return -1.0;
            }
        }
        *dt = dtm;
        return tmin;
    }

    static long long eval(Data *d, const double *u, long long dims, double z, double *f,
                          double *J, double *T, double *rates)
    {
        if (!d || !u || !f) return set_error("null argument");
        if (dims <= 0) return 0;
        DevState *s = state(d);
        if (ensure_device(d, s)) return -1;
        const size_t bu = (size_t)dims * N * sizeof(double);
        double *du = nullptr, *df = nullptr, *dJ = nullptr, *dT = nullptr, *dr = nullptr;
        int *dn = nullptr;
        DENGO_HIP_OK(hipMalloc(&du, bu));
        DENGO_HIP_OK(hipMalloc(&df, bu));
        DENGO_HIP_OK(hipMalloc(&dn, dims * sizeof(int)));
        if (J) DENGO_HIP_OK(hipMalloc(&dJ, bu * N));
        if (T) DENGO_HIP_OK(hipMalloc(&dT, dims * sizeof(double)));
        if (rates) DENGO_HIP_OK(hipMalloc(&dr, (size_t)dims * NTAB * sizeof(double)));
        DENGO_HIP_OK(hipMemcpy(du, u, bu, hipMemcpyHostToDevice));
        // a cell whose RHS sees a negative value gets neither f nor J: both
        // stay zero
        DENGO_HIP_OK(hipMemset(df, 0, bu));
        if (dJ) DENGO_HIP_OK(hipMemset(dJ, 0, bu * N));
        const long long blocks = (dims + 63) / 64;
        hipLaunchKernelGGL(k_eval<Net>, dim3((unsigned)blocks), dim3(64), 0, 0, du, dims,
                           env(d, s, z), df, dJ, dT, dr, dn);
        DENGO_HIP_OK(hipGetLastError());
        std::vector<int> neg(dims);
        DENGO_HIP_OK(hipMemcpy(neg.data(), dn, dims * sizeof(int), hipMemcpyDeviceToHost));
        DENGO_HIP_OK(hipMemcpy(f, df, bu, hipMemcpyDeviceToHost));
        if (J) DENGO_HIP_OK(hipMemcpy(J, dJ, bu * N, hipMemcpyDeviceToHost));
        if (T) DENGO_HIP_OK(hipMemcpy(T, dT, dims * sizeof(double), hipMemcpyDeviceToHost));
        if (rates)
            DENGO_HIP_OK(hipMemcpy(rates, dr, (size_t)dims * NTAB * sizeof(double),
                                   hipMemcpyDeviceToHost));
        void *ptrs[] = {du, df, dJ, dT, dr, dn};
        for (void *q : ptrs)
            if (q) DENGO_HIP_OK(hipFree(q));
        long long nneg = 0;
        for (long long i = 0; i < dims; ++i) nneg += neg[i];
        return nneg;
    }

    // BE_chem_solve(calculate_rhs, calculate_jacobian, u, dt, rtol, atol, nstrip,
    // nchem, scaling, data, ...) (BE_chem_solve.C:79-241) for host arrays: the
    // strip's cells share dt, and the Newton loop runs until every cell passes
    // the max-norm test (at most 10 sweeps).  Each sweep's per-cell work runs
    // on the GPU; the host keeps the strip state and applies the reference's
    // strip-wide outcomes in its order: a negative species in any cell's RHS
    // (BE_Resid_Fun, :246-275) or Jacobian (BE_Resid_Jac, :280-321) ends the
    // call with every cell's state left multiplied by the scaling at that
    // point; a NaN in cell x ends it after cells 0..x are updated; the
    // scaling loops run on the host in the reference's order.
    static int be_chem_solve(Data *d, double *u, double dt, double *rtol, double *atol,
                             int nstrip, int nchem, double *scaling)
    {
        if (!d || !u || !rtol || !atol || !scaling) return set_error("null argument");
        if (nchem != N) return set_error("nchem must be %d", N);
        if (nstrip <= 0) return 0;
        DevState *s = state(d);
        if (ensure_device(d, s) || ensure_scratch(s, nstrip)) return -1;
        if (nstrip > s->strip_cap) {
            if (grow(&s->d_u0, (long long)nstrip * N) || grow(&s->d_sc, (long long)nstrip * N) ||
                grow(&s->d_inv, (long long)nstrip * N) || grow(&s->d_unew, (long long)nstrip * N) ||
                grow(&s->d_code, (long long)nstrip))
                return -1;
            s->strip_cap = nstrip;
        }
        const size_t m = (size_t)nstrip * N;
        std::vector<double> inv(m), unew(m);
        std::vector<int> code(nstrip);
        for (size_t i = 0; i < m; ++i) inv[i] = 1.0 / scaling[i];
        for (size_t i = 0; i < m; ++i) u[i] *= inv[i];
        for (size_t i = 0; i < m; ++i) atol[i] *= inv[i];
        const size_t b = m * sizeof(double);
        int rc = 0;
        if (hipMemcpy(s->d_u0, u, b, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s->d_sc, scaling, b, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s->d_inv, inv.data(), b, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s->d_atol, atol, b, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s->d_rtol, rtol, b, hipMemcpyHostToDevice) != hipSuccess)
            rc = set_error("host-to-device copy failed");
        const typename Net::Env e = env(d, s, d->current_z);
        const unsigned blocks = (unsigned)((nstrip + 63) / 64);
        int rv = 1;
        bool ended = false;
        for (int sweep = 0; sweep < 10 && rc == 0 && !ended; ++sweep) {
            if (hipMemcpy(s->d_in, u, b, hipMemcpyHostToDevice) != hipSuccess) {
                rc = set_error("host-to-device copy failed");
                break;
            }
            hipLaunchKernelGGL(k_strip_sweep<Net>, dim3(blocks), dim3(64), 0, 0, s->d_in,
                               s->d_unew, s->d_u0, s->d_sc, s->d_inv, s->d_atol, s->d_rtol, dt,
                               (long long)nstrip, e, s->d_T, s->d_code);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpy(unew.data(), s->d_unew, b, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(code.data(), s->d_code, nstrip * sizeof(int), hipMemcpyDeviceToHost) !=
                    hipSuccess) {
                rc = set_error("strip sweep failed");
                break;
            }
            bool rhs_fail = false, jac_fail = false, unsolved = false;
            int first_nan = -1;
            for (int x = 0; x < nstrip; ++x) {
                rhs_fail = rhs_fail || code[x] == 2 + FK_RHS;
                jac_fail = jac_fail || code[x] == 2 + FK_JAC;
                if (first_nan < 0 && code[x] == 2 + FK_NAN) first_nan = x;
                unsolved = unsolved || code[x] == 1;
            }
            if (rhs_fail) {
                for (size_t i = 0; i < m; ++i) u[i] *= scaling[i];
                ended = true;
            } else if (jac_fail) {
                for (size_t i = 0; i < m; ++i) u[i] = ((u[i] * scaling[i]) * inv[i]) * scaling[i];
                ended = true;
            } else if (first_nan >= 0) {
                // cells 0..first_nan were updated; the rest keep the state after
                // the residual's and the Jacobian's scaling round trips
                for (size_t i = 0; i < m; ++i) {
                    if ((long long)(i / N) <= first_nan) u[i] = unew[i];
                    else u[i] = (((u[i] * scaling[i]) * inv[i]) * scaling[i]) * inv[i];
                }
                ended = true;
            } else {
                memcpy(u, unew.data(), b);
                if (!unsolved) {
                    rv = 0;
                    ended = true;
                }
            }
        }
        const int nts = nstrip < MAX_NCELLS ? nstrip : MAX_NCELLS;
        if (rc == 0 && hipMemcpy(d->Ts, s->d_T, nts * sizeof(double), hipMemcpyDeviceToHost) !=
                           hipSuccess)
            rc = set_error("device-to-host copy failed");
        for (size_t i = 0; i < m; ++i) u[i] *= scaling[i];
        for (size_t i = 0; i < m; ++i) atol[i] *= scaling[i];
        if (rc) return rc;
        return rv;
    }

    // temperature_from_mass_density (rates_and_rate_tables.C.template:635-668)
    // for nstrip cells of host memory, computed on the device.
    static int temperature_from_mass(const double *input, int nstrip, int nchem, double *T)
    {
        if (!input || !T) return set_error("null argument");
        if (nchem != N) return set_error("nchem must be %d", N);
        if (nstrip <= 0) return 0;
        double *du = nullptr, *dT = nullptr;
        const size_t bu = (size_t)nstrip * N * sizeof(double), bT = (size_t)nstrip * sizeof(double);
        int rc = 0;
        if (hipMalloc(&du, bu) != hipSuccess || hipMalloc(&dT, bT) != hipSuccess)
            rc = set_error("hipMalloc failed");
        if (rc == 0 && hipMemcpy(du, input, bu, hipMemcpyHostToDevice) != hipSuccess)
            rc = set_error("host-to-device copy failed");
        if (rc == 0) {
            hipLaunchKernelGGL(k_temperature_from_mass<Net>, dim3((unsigned)((nstrip + 63) / 64)),
                               dim3(64), 0, 0, du, (long long)nstrip, dT);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpy(T, dT, bT, hipMemcpyDeviceToHost) != hipSuccess)
                rc = set_error("temperature kernel failed");
        }
        if (du) (void)hipFree(du);
        if (dT) (void)hipFree(dT);
        return rc;
    }

    // ensure_electron_consistency (rates_and_rate_tables.C.template:604-630):
    // de of cell i is the running total of free electrons over cells 0..i of
    // the strip (the reference does not reset its total per cell).
    static void electron_consistency(double *input, int nstrip, int nchem)
    {
        if (!input || nchem != N) return;
        double total_e = 0.0;
        for (int i = 0; i < nstrip; ++i) {
            double *c = input + (size_t)i * N;
            for (int k = 0; k < N; ++k)
                if (k != Net::GE && k != Net::DE) total_e += c[Net::ext(k)] * Net::free_e(k);
            c[Net::ext(Net::DE)] = total_e;
        }
    }

    // Hooks of the patch server (dengo_server.hip).
    static int hook_prepare(void *data, const dengo_options *o, int *full)
    {
        Data *d = (Data *)data;
        dengo_options oo;
        if (check_options(o, &oo)) return -1;
        DevState *s = state(d);
        if (ensure_device(d, s)) return -1;
        return full_grid(s, oo, full);
    }

    static int hook_resident(void *data, const dengo_options *o, int *full)
    {
        Data *d = (Data *)data;
        dengo_options oo;
        if (check_options(o, &oo)) return -1;
        int dev = 0;
        DENGO_HIP_OK(hipGetDevice(&dev));
        DevState *s = (DevState *)d->device_state;
        if (!s || s->device != dev || s->stale || !s->pairs || !s->d_counter)
            return set_error("the tables of this data struct are not resident on device %d (one "
                             "evolve or eval call on that device uploads them)", dev);
        return full_grid(s, oo, full);
    }

    static int hook_prepare_order(void *data, OrderBufs *ob, const long long *order, long long n,
                                  void *stream)
    {
        Data *d = (Data *)data;
        DevState *s = state(d);
        if (ensure_device(d, s)) return -1;
        if (!order || n <= 0) return 0;
        OrderBufs &b = ob ? *ob : s->ord;
        return order_alloc(b, n, N) || order_check(b, order, n, (hipStream_t)stream) ? -1 : 0;
    }

    static int hook_launch(void *data, const LaunchArgs &a)
    {
        Data *d = (Data *)data;
        dengo_options o;
        if (check_options(&a.opt, &o)) return -1;
        return launch(d, state(d), o, a.dtf, a.dt0, a.z, a.x, a.rtol, a.atol, a.n, a.out,
                      a.stream, a.counter, false, a.ord);
    }

    static NetHooks hooks()
    {
        NetHooks h;
        h.nspecies = N;
        h.isothermal = 0;
        h.prepare = hook_prepare;
        h.prepare_order = hook_prepare_order;
        h.launch = hook_launch;
        h.resident = hook_resident;
        return h;
    }

    static int calculate_rhs(double *input, double *rhs, int nstrip, int nchem, void *sdata)
    {
        Data *d = (Data *)sdata;
        if (nchem != N) return 1;
        std::vector<double> T(nstrip);
        const long long r = eval(d, input, nstrip, d->current_z, rhs, NULL, T.data(), NULL);
        if (r < 0) {
            fprintf(stderr, "calculate_rhs: %s\n", last_error_text());
            return 1;
        }
        for (int i = 0; i < nstrip && i < MAX_NCELLS; ++i) d->Ts[i] = T[i];
        return r > 0 ? 1 : 0;
    }

    static int calculate_jacobian(double *input, double *Joutput, int nstrip, int nchem,
                                  void *sdata)
    {
        Data *d = (Data *)sdata;
        if (nchem != N) return 1;
        std::vector<double> f((size_t)nstrip * nchem);
        const long long r = eval(d, input, nstrip, d->current_z, f.data(), Joutput, NULL, NULL);
        if (r < 0) {
            fprintf(stderr, "calculate_jacobian: %s\n", last_error_text());
            return 1;
        }
        return r > 0 ? 1 : 0;
    }

    static void calculate_temperature(Data *d, double *input, int nstrip, int nchem)
    {
        if (nchem != N || nstrip <= 0) return;
        std::vector<double> f((size_t)nstrip * nchem), T(nstrip);
        if (eval(d, input, nstrip, d->current_z, f.data(), NULL, T.data(), NULL) < 0) {
            fprintf(stderr, "calculate_temperature: %s\n", last_error_text());
            return;
        }
        for (int i = 0; i < nstrip && i < MAX_NCELLS; ++i) d->Ts[i] = T[i];
    }
};

}  // namespace dengo

// extern "C" wrappers of one network: <n>_setup_data, dengo_evolve_<n>, ...
#define DENGO_DEFINE_C_ABI(n, H)                                                               \
    extern "C" {                                                                               \
    n##_data *n##_alloc_data(void) { return H::alloc(); }                                      \
    int n##_set_tables(n##_data *d, const double *t, int nt, int nT, double T0, double T1)     \
    {                                                                                          \
        return H::set_tables(d, t, nt, nT, T0, T1);                                            \
    }                                                                                          \
    n##_data *n##_setup_data_from(const char *path, int *nf, char ***fn)                       \
    {                                                                                          \
        return H::setup_from(path, nf, fn);                                                    \
    }                                                                                          \
    n##_data *n##_setup_data(int *nf, char ***fn)                                              \
    {                                                                                          \
        char b_[512];                                                                          \
        return H::setup_from(H::default_file(#n, "tables", b_, sizeof(b_)), nf, fn);           \
    }                                                                                          \
    void n##_read_rate_tables(n##_data *d)                                                     \
    {                                                                                          \
        char b_[512];                                                                          \
        if (H::read_tables(H::default_file(#n, "tables", b_, sizeof(b_)), d) != 0)             \
            fprintf(stderr, "%s\n", dengo_last_error());                                       \
    }                                                                                          \
    void n##_read_cooling_tables(n##_data *d)                                                  \
    {                                                                                          \
        char b_[512];                                                                          \
        if (H::read_tables(H::default_file(#n, "tables", b_, sizeof(b_)), d) != 0)             \
            fprintf(stderr, "%s\n", dengo_last_error());                                       \
    }                                                                                          \
    int n##_main(int argc, char **argv)                                                        \
    {                                                                                          \
        (void)argc;                                                                            \
        (void)argv;                                                                            \
        return H::main_driver(#n);                                                             \
    }                                                                                          \
    size_t n##_data_bytes(void) { return sizeof(n##_data); }                                   \
    void n##_free_data(n##_data *d) { H::free_data(d); }                                       \
    int dengo_##n##_be_chem_solve(n##_data *d, double *u, double dt, double *rt, double *at,   \
                                  int nstrip, int nchem, double *scaling)                      \
    {                                                                                          \
        return H::be_chem_solve(d, u, dt, rt, at, nstrip, nchem, scaling);                     \
    }                                                                                          \
    void ensure_electron_consistency_##n(double *input, int nstrip, int nchem)                 \
    {                                                                                          \
        H::electron_consistency(input, nstrip, nchem);                                         \
    }                                                                                          \
    void temperature_from_mass_density_##n(double *input, int nstrip, int nchem, double *T)    \
    {                                                                                          \
        if (H::temperature_from_mass(input, nstrip, nchem, T) != 0)                            \
            fprintf(stderr, "temperature_from_mass_density_" #n ": %s\n", dengo_last_error()); \
    }                                                                                          \
    void dengo_##n##_set_current_z(n##_data *d, double z)                                      \
    {                                                                                          \
        if (d) d->current_z = z;                                                               \
    }                                                                                          \
    int dengo_##n##_strip_temperatures(const n##_data *d, double *T, int nstrip)               \
    {                                                                                          \
        if (!d || !T) return ::dengo::set_error("null argument");                              \
        const int m = nstrip < MAX_NCELLS ? nstrip : MAX_NCELLS;                               \
        for (int i = 0; i < m; ++i) T[i] = d->Ts[i];                                           \
        return m;                                                                              \
    }                                                                                          \
    int dengo_##n##_evolve_device(n##_data *d, const dengo_options *o, double dtf, double dt0,  \
                                  double z, double *x, const double *rt, double *at,           \
                                  long long dims, const dengo_cell_outputs *out,              \
                                  double *tmin, double *dtmin, void *stream)                  \
    {                                                                                          \
        return H::evolve_device(d, o, dtf, dt0, z, x, rt, at, dims, out, tmin, dtmin, stream); \
    }                                                                                          \
    double dengo_##n##_evolve_host(n##_data *d, const dengo_options *o, double dtf, double *dt, \
                                   double z, double *x, double *rt, double *at, long long dims, \
                                   int *st, int *ns, int *nr, int *nw, double *dtc, double *Tc) \
    {                                                                                          \
        return H::evolve_host(d, o, dtf, dt, z, x, rt, at, dims, st, ns, nr, nw, dtc, Tc);     \
    }                                                                                          \
    const dengo_net_hooks *dengo_##n##_server_hooks(void)                                      \
    {                                                                                          \
        static const dengo::NetHooks h = H::hooks();                                           \
        return reinterpret_cast<const dengo_net_hooks *>(&h);                                  \
    }                                                                                          \
    int dengo_##n##_set_write_intermediate(n##_data *d, int on)                                \
    {                                                                                          \
        return H::set_write_intermediate(d, on);                                               \
    }                                                                                          \
    double dengo_evolve_##n(double dtf, double *dt, double z, double *input, double *rtol,     \
                            double *atol, long long dims, n##_data *data)                      \
    {                                                                                          \
        dengo_options o;                                                                       \
        dengo_default_options(&o);                                                             \
        const double t = H::evolve_host(data, &o, dtf, dt, z, input, rtol, atol, dims, NULL,   \
                                        NULL, NULL, NULL, NULL, NULL);                         \
        if (t < 0) fprintf(stderr, "dengo_evolve_" #n ": %s\n", dengo_last_error());           \
        return t;                                                                              \
    }                                                                                          \
    long long dengo_##n##_eval(n##_data *d, const double *u, long long dims, double z,         \
                               double *f, double *J, double *T, double *rates)                 \
    {                                                                                          \
        return H::eval(d, u, dims, z, f, J, T, rates);                                         \
    }                                                                                          \
    int calculate_rhs_##n(double *in, double *rhs, int ns, int nc, void *sd)                   \
    {                                                                                          \
        return H::calculate_rhs(in, rhs, ns, nc, sd);                                          \
    }                                                                                          \
    int calculate_jacobian_##n(double *in, double *J, int ns, int nc, void *sd)                \
    {                                                                                          \
        return H::calculate_jacobian(in, J, ns, nc, sd);                                       \
    }                                                                                          \
    void n##_calculate_temperature(n##_data *d, double *in, int ns, int nc)                    \
    {                                                                                          \
        H::calculate_temperature(d, in, ns, nc);                                               \
    }                                                                                          \
    }

// The reference's unsuffixed helpers (rates_and_rate_tables.h.template:106-107),
// exported only by a library that holds a single network (the solver
// libraries write_solver emits): their names carry no network.
#define DENGO_DEFINE_SINGLE_NETWORK_ABI(n)                                                     \
    extern "C" {                                                                               \
    void ensure_electron_consistency(double *input, int nstrip, int nchem)                     \
    {                                                                                          \
        ensure_electron_consistency_##n(input, nstrip, nchem);                                 \
    }                                                                                          \
    void temperature_from_mass_density(double *input, int nstrip, int nchem, double *T)        \
    {                                                                                          \
        temperature_from_mass_density_##n(input, nstrip, nchem, T);                            \
    }                                                                                          \
    }
