// Primordial-network kernels and the C ABI of the HIP backend.
//
// Kernels (one cell per lane, state in registers, tables in HBM/L2):
//   k_prim<false, MODE>  one cell per lane, grid covers all cells
//   k_prim<true, MODE>   persistent: resident lanes pull cells from a device
//                        counter, wave-aggregated, one Newton sweep per loop
//   MODE: kFaithful (reference-faithful FP64), kFast64 (block-structured FP64
//         solve), kMixed (FP32 LU of the equilibrated block, FP64 residual)
//   k_prim_eval          batched RHS / Jacobian / temperature evaluation
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "be_driver.h"
#include "be_fast.h"
#include "dengo_hip.h"

using namespace dengo;

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
constexpr int kBlock = 64;

enum { kFaithful = 0, kFast64 = 1, kMixed = 2 };

template <int MODE>
DENGO_D bool cell_start(Cell<PrimNet> &c, long long id, const KParams<PrimNet> &p)
{
    if constexpr (MODE == kFaithful) return start_or_finish(c, id, p);
    else return start_or_finish_fast(c, id, p);
}

template <int MODE>
DENGO_D bool cell_sweep(Cell<PrimNet> &c, const KParams<PrimNet> &p)
{
    if constexpr (MODE == kFaithful) {
        const int code = sweep_fp64(c, p);
        return code != SW_CONTINUE && end_be(c, code, p);
    } else {
        const int code = sweep_fast<PrimNet, typename std::conditional<MODE == kMixed, float,
                                                                       double>::type>(c, p);
        return code != SW_CONTINUE && end_be_fast(c, code, p);
    }
}

template <bool PERSIST, int MODE>
__global__ void __launch_bounds__(kBlock) k_prim(KParams<PrimNet> p)
{
    Cell<PrimNet> c;
    bool active = false;
    bool exhausted = false;
    if (!PERSIST) {
        const long long id = (long long)blockIdx.x * kBlock + threadIdx.x;
        if (id < p.ncells) active = cell_start<MODE>(c, id, p);
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
                    if (id < (unsigned long long)p.ncells) active = cell_start<MODE>(c, (long long)id, p);
                    else exhausted = true;
                }
            }
            if (!__any(active || !exhausted)) break;
        } else {
            if (!__any(active)) break;
        }
        if (active && cell_sweep<MODE>(c, p)) {
            finish_cell(c, p);
            active = false;
        }
    }
}

__global__ void k_reset_counter(unsigned long long *counter) { *counter = 0ull; }

// Smallest elapsed time over cells (ties: lowest index) and its final dt.
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

__global__ void __launch_bounds__(64) k_prim_eval(const double *u_in, long long n, prim::Env env,
                                                  double *f_out, double *J_out, double *T_out,
                                                  double *rates_out, int *neg_out)
{
    const long long id = (long long)blockIdx.x * 64 + threadIdx.x;
    if (id >= n) return;
    double u[prim::N], f[prim::N], rs[prim::NTAB], J[prim::N * prim::N];
    double T;
#pragma unroll
    for (int k = 0; k < prim::N; ++k) u[k] = u_in[id * prim::N + k];
    const bool ok = prim::rhs(u, f, rs, T, env);
    if (T_out) T_out[id] = T;
    if (rates_out)
#pragma unroll
        for (int t = 0; t < prim::NTAB; ++t) rates_out[id * prim::NTAB + t] = rs[t];
    neg_out[id] = ok ? 0 : 1;
    if (!ok) return;
#pragma unroll
    for (int k = 0; k < prim::N; ++k) f_out[id * prim::N + k] = f[k];
    if (J_out) {
        prim::jacobian(u, rs, T, env, J);
#pragma unroll
        for (int k = 0; k < prim::N * prim::N; ++k) J_out[id * prim::N * prim::N + k] = J[k];
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace {

thread_local char g_err[512] = "";
thread_local double g_last_ms = 0.0;

int fail(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return -1;
}

#define HIP_OK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                     \
    } while (0)

struct DevState {
    int device = -1;
    double2 *pairs = nullptr;
    int nbins = 0;
    bool stale = true;
    // scratch owned by the library
    long long cap = 0;
    double *d_in = nullptr, *d_rtol = nullptr, *d_atol = nullptr;
    double *d_ttot = nullptr, *d_dt = nullptr, *d_T = nullptr;
    int *d_status = nullptr, *d_nsub = nullptr, *d_nrej = nullptr, *d_nsweep = nullptr;
    int *d_neg = nullptr;
    unsigned long long *d_counter = nullptr;
    double *d_red = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int cu_count = 0;
};

DevState *state_of(primordial_data *d)
{
    if (!d->device_state) d->device_state = new DevState();
    return (DevState *)d->device_state;
}

const double *table_ptr(const primordial_data *d, int t)
{
    static const size_t offs[] = {
#define X(f) offsetof(primordial_data, f),
        PRIMORDIAL_TABLE_FIELDS(X)
#undef X
    };
    return (const double *)((const char *)d + offs[t]);
}

double *table_ptr_mut(primordial_data *d, int t)
{
    return (double *)table_ptr(d, t);
}

void set_grid(primordial_data *d, double T0, double T1)
{
    // rates_and_rate_tables.C.template:25-36
    d->bounds[0] = T0;
    d->bounds[1] = T1;
    d->nbins = PRIMORDIAL_NT - 1;
    d->dbin = (log(d->bounds[1]) - log(d->bounds[0])) / d->nbins;
    d->idbin = (double)(1.0L / d->dbin);
    d->z_bounds[0] = 0.0;
    d->z_bounds[1] = 0.0;
    d->n_zbins = 0;
    d->d_zbin = 0.0;
    d->id_zbin = 0.0;
}

int ensure_device(primordial_data *d, DevState *s)
{
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (s->device != dev) {
        if (s->device >= 0) return fail("data struct was set up on device %d, current is %d",
                                        s->device, dev);
        s->device = dev;
        HIP_OK(hipMalloc(&s->d_counter, sizeof(unsigned long long)));
        HIP_OK(hipMalloc(&s->d_red, 2 * sizeof(double)));
        HIP_OK(hipEventCreate(&s->ev0));
        HIP_OK(hipEventCreate(&s->ev1));
        hipDeviceProp_t prop;
        HIP_OK(hipGetDeviceProperties(&prop, dev));
        s->cu_count = prop.multiProcessorCount;
    }
    if (s->stale || !s->pairs) {
        if (d->nbins != PRIMORDIAL_NT - 1) return fail("tables not set (nbins=%d)", d->nbins);
        const int nb = d->nbins;
        std::vector<double2> host((size_t)nb * PRIMORDIAL_NTABLES);
        for (int t = 0; t < PRIMORDIAL_NTABLES; ++t) {
            const double *r = table_ptr(d, t);
            for (int b = 0; b < nb; ++b) host[(size_t)b * PRIMORDIAL_NTABLES + t] = make_double2(r[b], r[b + 1]);
        }
        if (!s->pairs) HIP_OK(hipMalloc(&s->pairs, host.size() * sizeof(double2)));
        HIP_OK(hipMemcpy(s->pairs, host.data(), host.size() * sizeof(double2),
                         hipMemcpyHostToDevice));
        s->nbins = nb;
        s->stale = false;
    }
    return 0;
}

template <class T>
int grow(T **p, long long n)
{
    if (*p) HIP_OK(hipFree(*p));
    *p = nullptr;
    HIP_OK(hipMalloc(p, (size_t)n * sizeof(T)));
    return 0;
}

int ensure_scratch(DevState *s, long long n)
{
    if (n <= s->cap) return 0;
    long long c = n;
    if (grow(&s->d_in, c * 10) || grow(&s->d_rtol, c * 10) || grow(&s->d_atol, c * 10) ||
        grow(&s->d_ttot, c) || grow(&s->d_dt, c) || grow(&s->d_T, c) || grow(&s->d_status, c) ||
        grow(&s->d_nsub, c) || grow(&s->d_nrej, c) || grow(&s->d_nsweep, c) || grow(&s->d_neg, c))
        return -1;
    s->cap = c;
    return 0;
}

prim::Env make_env(const primordial_data *d, const DevState *s, double z)
{
    prim::Env env;
    env.tb.pairs = s->pairs;
    env.tb.nbins = d->nbins;
    env.tb.dbin = d->dbin;
    env.tb.idbin = d->idbin;
    env.tb.lb = log(d->bounds[0]);
    env.tb.T0 = d->bounds[0];
    env.tb.T1 = d->bounds[1];
    env.zp1 = z + 1.0;
    env.comp4 = pow(env.zp1, 4);
    return env;
}

int read_blob(const char *path, primordial_data *d)
{
    FILE *f = fopen(path, "rb");
    if (!f) return fail("cannot open table file %s", path);
    char magic[8];
    int32_t nt = 0, nT = 0;
    double hdr[4];
    char pad[64 - 8 - 8 - 32];
    int ok = fread(magic, 1, 8, f) == 8 && fread(&nt, 4, 1, f) == 1 && fread(&nT, 4, 1, f) == 1 &&
             fread(hdr, 8, 4, f) == 4 && fread(pad, 1, sizeof(pad), f) == sizeof(pad);
    if (!ok || memcmp(magic, "DENGOTB1", 8) != 0) {
        fclose(f);
        return fail("%s is not a DENGOTB1 table blob", path);
    }
    if (nt != PRIMORDIAL_NTABLES || nT != PRIMORDIAL_NT) {
        fclose(f);
        return fail("%s holds %d tables x %d bins, the primordial struct needs %d x %d", path,
                    nt, nT, PRIMORDIAL_NTABLES, PRIMORDIAL_NT);
    }
    for (int t = 0; t < nt; ++t) {
        if (fread(table_ptr_mut(d, t), 8, nT, f) != (size_t)nT) {
            fclose(f);
            return fail("%s: short read", path);
        }
    }
    fclose(f);
    set_grid(d, hdr[0], hdr[1]);
    state_of(d)->stale = true;
    return 0;
}

const char *kFieldNames[PRIMORDIAL_NSPECIES] = {"H2I", "H2II", "HI", "HII", "HM",
                                                "HeI", "HeII", "HeIII", "de", "ge"};

void field_names(int *NumberOfFields, char ***FieldNames)
{
    if (FieldNames != NULL && NumberOfFields != NULL) {
        NumberOfFields[0] = PRIMORDIAL_NSPECIES;
        FieldNames[0] = new char *[PRIMORDIAL_NSPECIES];
        for (int i = 0; i < PRIMORDIAL_NSPECIES; ++i) FieldNames[0][i] = strdup(kFieldNames[i]);
    }
}

int launch_evolve(primordial_data *d, DevState *s, const dengo_options &o, double dtf, double dt0,
                  double z, double *d_input, const double *d_rtol, double *d_atol, long long dims,
                  const dengo_cell_outputs &out, hipStream_t stream)
{
    KParams<PrimNet> p;
    p.input = d_input;
    p.rtol = d_rtol;
    p.atol = d_atol;
    p.ncells = dims;
    p.env = make_env(d, s, z);
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
    p.counter = s->d_counter;
    if (dims <= 0) return 0;
    HIP_OK(hipEventRecord(s->ev0, stream));
    const int mode = o.precision == DENGO_FP64 ? kFaithful
                   : o.precision == DENGO_FP64_FAST ? kFast64 : kMixed;
    void (*kp)(KParams<PrimNet>) = nullptr;
    void (*kc)(KParams<PrimNet>) = nullptr;
    if (mode == kFaithful) { kp = k_prim<true, kFaithful>; kc = k_prim<false, kFaithful>; }
    else if (mode == kFast64) { kp = k_prim<true, kFast64>; kc = k_prim<false, kFast64>; }
    else { kp = k_prim<true, kMixed>; kc = k_prim<false, kMixed>; }
    if (o.schedule == DENGO_PERSISTENT) {
        int grid = o.grid;
        if (grid <= 0) {
            int per_cu = 0;
            HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kp, kBlock, 0));
            if (per_cu < 1) per_cu = 1;
            grid = per_cu * s->cu_count;
        }
        const long long need = (dims + kBlock - 1) / kBlock;
        if ((long long)grid > need) grid = (int)need;
        hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, stream, s->d_counter);
        HIP_OK(hipGetLastError());
        hipLaunchKernelGGL(kp, dim3(grid), dim3(kBlock), 0, stream, p);
    } else {
        const long long blocks = (dims + kBlock - 1) / kBlock;
        if (blocks > 0x7fffffffLL) return fail("too many cells: %lld", dims);
        hipLaunchKernelGGL(kc, dim3((unsigned)blocks), dim3(kBlock), 0, stream, p);
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(s->ev1, stream));
    return 0;
}

int check_opts(const dengo_options *opt, dengo_options *o)
{
    dengo_default_options(o);
    if (opt) *o = *opt;
    if (o->block != 0 && o->block != kBlock) return fail("block must be %d", kBlock);
    if (o->max_substeps <= 0) o->max_substeps = 50000;
    if (o->schedule != DENGO_PER_CELL && o->schedule != DENGO_PERSISTENT)
        return fail("unknown schedule %d", o->schedule);
    if (o->precision != DENGO_FP64 && o->precision != DENGO_MIXED &&
        o->precision != DENGO_FP64_FAST)
        return fail("unknown precision %d", o->precision);
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

void dengo_default_options(dengo_options *opt)
{
    memset(opt, 0, sizeof(*opt));
    opt->precision = DENGO_FP64;
    opt->schedule = DENGO_PERSISTENT;
    opt->block = kBlock;
    opt->max_substeps = 50000;
}

const char *dengo_last_error(void) { return g_err; }

double dengo_last_kernel_ms(void) { return g_last_ms; }

primordial_data *primordial_alloc_data(void)
{
    primordial_data *d = (primordial_data *)calloc(1, sizeof(primordial_data));
    return d;
}

int primordial_set_tables(primordial_data *d, const double *tables, int ntables, int nT,
                          double T0, double T1)
{
    if (!d || !tables) return fail("null argument");
    if (ntables != PRIMORDIAL_NTABLES || nT != PRIMORDIAL_NT)
        return fail("expected %d tables x %d bins, got %d x %d", PRIMORDIAL_NTABLES,
                    PRIMORDIAL_NT, ntables, nT);
    if (!(T0 > 0) || !(T1 > T0)) return fail("bad temperature bounds");
    for (int t = 0; t < ntables; ++t)
        memcpy(table_ptr_mut(d, t), tables + (size_t)t * nT, sizeof(double) * nT);
    set_grid(d, T0, T1);
    state_of(d)->stale = true;
    return 0;
}

primordial_data *primordial_setup_data_from(const char *path, int *NumberOfFields,
                                            char ***FieldNames)
{
    primordial_data *d = primordial_alloc_data();
    if (!d) return NULL;
    if (read_blob(path, d) != 0) {
        fprintf(stderr, "%s\n", g_err);
        free(d);
        return NULL;
    }
    field_names(NumberOfFields, FieldNames);
    return d;
}

primordial_data *primordial_setup_data(int *NumberOfFields, char ***FieldNames)
{
    return primordial_setup_data_from("primordial_tables.bin", NumberOfFields, FieldNames);
}

void primordial_read_rate_tables(primordial_data *d)
{
    if (read_blob("primordial_tables.bin", d) != 0) fprintf(stderr, "%s\n", g_err);
}

void primordial_read_cooling_tables(primordial_data *d)
{
    if (read_blob("primordial_tables.bin", d) != 0) fprintf(stderr, "%s\n", g_err);
}

void primordial_free_data(primordial_data *d)
{
    if (!d) return;
    DevState *s = (DevState *)d->device_state;
    if (s) {
        void *ptrs[] = {s->pairs, s->d_in, s->d_rtol, s->d_atol, s->d_ttot, s->d_dt, s->d_T,
                        s->d_status, s->d_nsub, s->d_nrej, s->d_nsweep, s->d_neg,
                        s->d_counter, s->d_red};
        for (void *q : ptrs)
            if (q) (void)hipFree(q);
        if (s->ev0) (void)hipEventDestroy(s->ev0);
        if (s->ev1) (void)hipEventDestroy(s->ev1);
        delete s;
    }
    free(d);
}

int dengo_primordial_evolve_device(primordial_data *d, const dengo_options *opt, double dtf,
                                   double dt0, double z, double *d_input, const double *d_rtol,
                                   double *d_atol, long long dims,
                                   const dengo_cell_outputs *outputs, double *ttot_min,
                                   double *dt_min, void *stream)
{
    if (!d || !d_input || !d_rtol || !d_atol) return fail("null argument");
    if (dims < 0) return fail("negative cell count");
    dengo_options o;
    if (check_opts(opt, &o)) return -1;
    DevState *s = state_of(d);
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
    d->current_z = z;
    if (launch_evolve(d, s, o, dtf, dt0, z, d_input, d_rtol, d_atol, dims, out, st)) return -1;
    if (ttot_min) {
        if (dims > 0) {
            hipLaunchKernelGGL(k_min_ttot, dim3(1), dim3(1024), 0, st, out.ttot, out.dt, dims,
                               s->d_red);
            HIP_OK(hipGetLastError());
            double red[2];
            HIP_OK(hipMemcpyAsync(red, s->d_red, sizeof(red), hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
            *ttot_min = red[0];
            if (dt_min) *dt_min = red[1];
        } else {
            *ttot_min = 0.0;
            if (dt_min) *dt_min = dt0;
        }
        float ms = 0.f;
        if (dims > 0 && hipEventElapsedTime(&ms, s->ev0, s->ev1) == hipSuccess) g_last_ms = ms;
    }
    return 0;
}

double dengo_primordial_evolve_host(primordial_data *d, const dengo_options *opt, double dtf,
                                    double *dt, double z, double *input, double *rtol,
                                    double *atol, long long dims, int *status, int *nsub,
                                    int *nrej, int *nsweep, double *dt_cell, double *T_cell)
{
    if (!d || !dt || !input || !rtol || !atol) {
        fail("null argument");
        return -1.0;
    }
    if (dims <= 0) return 0.0;
    DevState *s = state_of(d);
    if (ensure_device(d, s) || ensure_scratch(s, dims)) return -1.0;
    const size_t bytes = (size_t)dims * PRIMORDIAL_NSPECIES * sizeof(double);
    if (hipMemcpy(s->d_in, input, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->d_rtol, rtol, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->d_atol, atol, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        fail("host-to-device copy failed");
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
    if (dengo_primordial_evolve_device(d, opt, dtf, *dt, z, s->d_in, s->d_rtol, s->d_atol, dims,
                                       &out, &tmin, &dtm, NULL) != 0)
        return -1.0;
#define D2H(dst, src, n)                                                                 \
    if ((dst) && hipMemcpy((dst), (src), (n), hipMemcpyDeviceToHost) != hipSuccess) {    \
        fail("device-to-host copy failed");                                              \
        return -1.0;                                                                     \
    }
    D2H(input, s->d_in, bytes);
    D2H(atol, s->d_atol, bytes);
    D2H(status, s->d_status, dims * sizeof(int));
    D2H(nsub, s->d_nsub, dims * sizeof(int));
    D2H(nrej, s->d_nrej, dims * sizeof(int));
    D2H(nsweep, s->d_nsweep, dims * sizeof(int));
    D2H(dt_cell, s->d_dt, dims * sizeof(double));
    D2H(T_cell, s->d_T, dims * sizeof(double));
    // the reference leaves the last strip's temperatures in data->Ts
    const long long nts = dims < MAX_NCELLS ? dims : MAX_NCELLS;
    D2H(d->Ts, s->d_T, nts * sizeof(double));
#undef D2H
    *dt = dtm;
    return tmin;
}

double dengo_evolve_primordial(double dtf, double *dt, double z, double *input, double *rtol,
                               double *atol, long long dims, primordial_data *data)
{
    dengo_options o;
    dengo_default_options(&o);
    const double t = dengo_primordial_evolve_host(data, &o, dtf, dt, z, input, rtol, atol, dims,
                                                  NULL, NULL, NULL, NULL, NULL, NULL);
    if (t < 0) fprintf(stderr, "dengo_evolve_primordial: %s\n", g_err);
    return t;
}

long long dengo_primordial_eval(primordial_data *d, const double *u, long long dims, double z,
                                double *f, double *J, double *T, double *rates)
{
    if (!d || !u || !f) return fail("null argument");
    if (dims <= 0) return 0;
    DevState *s = state_of(d);
    if (ensure_device(d, s)) return -1;
    const size_t bu = (size_t)dims * prim::N * sizeof(double);
    double *du = nullptr, *df = nullptr, *dJ = nullptr, *dT = nullptr, *dr = nullptr;
    int *dn = nullptr;
    HIP_OK(hipMalloc(&du, bu));
    HIP_OK(hipMalloc(&df, bu));
    HIP_OK(hipMalloc(&dn, dims * sizeof(int)));
    if (J) HIP_OK(hipMalloc(&dJ, bu * prim::N));
    if (T) HIP_OK(hipMalloc(&dT, dims * sizeof(double)));
    if (rates) HIP_OK(hipMalloc(&dr, (size_t)dims * prim::NTAB * sizeof(double)));
    HIP_OK(hipMemcpy(du, u, bu, hipMemcpyHostToDevice));
    HIP_OK(hipMemset(df, 0, bu));
    const long long blocks = (dims + 63) / 64;
    hipLaunchKernelGGL(k_prim_eval, dim3((unsigned)blocks), dim3(64), 0, 0, du, dims,
                       make_env(d, s, z), df, dJ, dT, dr, dn);
    HIP_OK(hipGetLastError());
    std::vector<int> neg(dims);
    HIP_OK(hipMemcpy(neg.data(), dn, dims * sizeof(int), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(f, df, bu, hipMemcpyDeviceToHost));
    if (J) HIP_OK(hipMemcpy(J, dJ, bu * prim::N, hipMemcpyDeviceToHost));
    if (T) HIP_OK(hipMemcpy(T, dT, dims * sizeof(double), hipMemcpyDeviceToHost));
    if (rates)
        HIP_OK(hipMemcpy(rates, dr, (size_t)dims * prim::NTAB * sizeof(double), hipMemcpyDeviceToHost));
    void *ptrs[] = {du, df, dJ, dT, dr, dn};
    for (void *q : ptrs)
        if (q) HIP_OK(hipFree(q));
    long long nneg = 0;
    for (long long i = 0; i < dims; ++i) nneg += neg[i];
    return nneg;
}

int calculate_rhs_primordial(double *input, double *rhs, int nstrip, int nchem, void *sdata)
{
    primordial_data *d = (primordial_data *)sdata;
    if (nchem != PRIMORDIAL_NSPECIES) return 1;
    std::vector<double> T(nstrip);
    const long long r = dengo_primordial_eval(d, input, nstrip, d->current_z, rhs, NULL, T.data(), NULL);
    if (r < 0) {
        fprintf(stderr, "calculate_rhs_primordial: %s\n", g_err);
        return 1;
    }
    for (int i = 0; i < nstrip && i < MAX_NCELLS; ++i) d->Ts[i] = T[i];
    return r > 0 ? 1 : 0;
}

int calculate_jacobian_primordial(double *input, double *Joutput, int nstrip, int nchem,
                                  void *sdata)
{
    primordial_data *d = (primordial_data *)sdata;
    if (nchem != PRIMORDIAL_NSPECIES) return 1;
    std::vector<double> f((size_t)nstrip * nchem);
    const long long r = dengo_primordial_eval(d, input, nstrip, d->current_z, f.data(), Joutput,
                                              NULL, NULL);
    if (r < 0) {
        fprintf(stderr, "calculate_jacobian_primordial: %s\n", g_err);
        return 1;
    }
    return r > 0 ? 1 : 0;
}

void primordial_calculate_temperature(primordial_data *d, double *input, int nstrip, int nchem)
{
    if (nchem != PRIMORDIAL_NSPECIES || nstrip <= 0) return;
    std::vector<double> f((size_t)nstrip * nchem), T(nstrip);
    if (dengo_primordial_eval(d, input, nstrip, d->current_z, f.data(), NULL, T.data(), NULL) < 0) {
        fprintf(stderr, "primordial_calculate_temperature: %s\n", g_err);
        return;
    }
    for (int i = 0; i < nstrip && i < MAX_NCELLS; ++i) d->Ts[i] = T[i];
}

}  // extern "C"
