// Host side and C ABI of the dense-Jacobian solver for the synthetic network.
#include <algorithm>
#include <numeric>
#include <vector>

#include "big_net.h"
#include "big_program.h"
#include "dengo_solver.h"
#include "synthetic_solver.h"

using namespace dengo;
namespace B = dengo::big;

static_assert(SYNTHETIC_NSPECIES == B::N, "network width");
static_assert(SYNTHETIC_MAX_REACTIONS == B::MAX_RXN, "reaction limit");

struct synthetic_data {
    int n = 0, nrxn = 0;
    std::vector<int> rxn;
    std::vector<double> A, beta, gamma;
    B::Programs prog;
    // device side
    int device = -1, cu_count = 0;
    int4 *d_rxn = nullptr;
    double *d_A = nullptr, *d_beta = nullptr, *d_gamma = nullptr;
    unsigned *d_rhs = nullptr, *d_jac = nullptr;
    unsigned long long *d_counter = nullptr;
    double *d_red = nullptr;
    long long cap = 0;
    double *d_x = nullptr, *d_T = nullptr, *d_rtol = nullptr, *d_atol = nullptr;
    double *d_dt = nullptr, *d_ttot = nullptr;
    int *d_status = nullptr, *d_nsub = nullptr, *d_nrej = nullptr, *d_nsweep = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    OrderBufs ord;                     // cell order check (the kernel reads the order itself)
};

namespace {

int ensure_device(synthetic_data *d)
{
    int dev = 0;
    DENGO_HIP_OK(hipGetDevice(&dev));
    if (d->device == dev) return 0;
    if (d->device >= 0)
        return set_error("data struct was set up on device %d, current is %d", d->device, dev);
    d->device = dev;
    DENGO_HIP_OK(hipDeviceGetAttribute(&d->cu_count, hipDeviceAttributeMultiprocessorCount, dev));
    std::vector<int4> r4(d->nrxn);
    for (int r = 0; r < d->nrxn; ++r)
        r4[r] = make_int4(d->rxn[4 * r], d->rxn[4 * r + 1], d->rxn[4 * r + 2], d->rxn[4 * r + 3]);
    DENGO_HIP_OK(hipMalloc(&d->d_rxn, sizeof(int4) * d->nrxn));
    DENGO_HIP_OK(hipMalloc(&d->d_A, sizeof(double) * d->nrxn));
    DENGO_HIP_OK(hipMalloc(&d->d_beta, sizeof(double) * d->nrxn));
    DENGO_HIP_OK(hipMalloc(&d->d_gamma, sizeof(double) * d->nrxn));
    DENGO_HIP_OK(hipMalloc(&d->d_rhs, sizeof(unsigned) * d->prog.rhs_ops.size()));
    DENGO_HIP_OK(hipMalloc(&d->d_jac, sizeof(unsigned) * d->prog.jac_ops.size()));
    DENGO_HIP_OK(hipMalloc(&d->d_counter, sizeof(unsigned long long)));
    DENGO_HIP_OK(hipMalloc(&d->d_red, 2 * sizeof(double)));
    DENGO_HIP_OK(hipMemcpy(d->d_rxn, r4.data(), sizeof(int4) * d->nrxn, hipMemcpyHostToDevice));
    DENGO_HIP_OK(hipMemcpy(d->d_A, d->A.data(), sizeof(double) * d->nrxn, hipMemcpyHostToDevice));
    DENGO_HIP_OK(hipMemcpy(d->d_beta, d->beta.data(), sizeof(double) * d->nrxn,
                           hipMemcpyHostToDevice));
    DENGO_HIP_OK(hipMemcpy(d->d_gamma, d->gamma.data(), sizeof(double) * d->nrxn,
                           hipMemcpyHostToDevice));
    DENGO_HIP_OK(hipMemcpy(d->d_rhs, d->prog.rhs_ops.data(), sizeof(unsigned) * d->prog.rhs_ops.size(),
                           hipMemcpyHostToDevice));
    DENGO_HIP_OK(hipMemcpy(d->d_jac, d->prog.jac_ops.data(), sizeof(unsigned) * d->prog.jac_ops.size(),
                           hipMemcpyHostToDevice));
    DENGO_HIP_OK(hipEventCreate(&d->ev0));
    DENGO_HIP_OK(hipEventCreate(&d->ev1));
    return 0;
}

B::NetDev net_view(const synthetic_data *d)
{
    B::NetDev v;
    v.nrxn = d->nrxn;
    v.rxn = d->d_rxn;
    v.A = d->d_A;
    v.beta = d->d_beta;
    v.gamma = d->d_gamma;
    v.rhs_ops = d->d_rhs;
    v.rhs_len = d->prog.rhs_len;
    v.jac_ops = d->d_jac;
    for (int J = 0; J < B::NT; ++J) {
        v.jac_len[J] = d->prog.jac_len[J];
        v.jac_off[J] = d->prog.jac_off[J];
    }
    for (int w = 0; w < B::NWAVE; ++w) v.own_off[w] = d->prog.own_off[w];
    v.own_cnt_off = d->prog.own_cnt_off;
    v.diag_len = d->prog.diag_len;
    v.diag_off = d->prog.diag_off;
    v.order = nullptr;
    return v;
}

int ensure_scratch(synthetic_data *d, long long dims)
{
    if (dims <= d->cap) return 0;
    double **dp[] = {&d->d_x, &d->d_T, &d->d_rtol, &d->d_atol, &d->d_dt, &d->d_ttot};
    int **ip[] = {&d->d_status, &d->d_nsub, &d->d_nrej, &d->d_nsweep};
    for (double **p : dp) {
        if (*p) DENGO_HIP_OK(hipFree(*p));
        *p = nullptr;
    }
    for (int **p : ip) {
        if (*p) DENGO_HIP_OK(hipFree(*p));
        *p = nullptr;
    }
    const size_t big = (size_t)dims * B::N * sizeof(double);
    DENGO_HIP_OK(hipMalloc(&d->d_x, big));
    DENGO_HIP_OK(hipMalloc(&d->d_rtol, big));
    DENGO_HIP_OK(hipMalloc(&d->d_atol, big));
    DENGO_HIP_OK(hipMalloc(&d->d_T, dims * sizeof(double)));
    DENGO_HIP_OK(hipMalloc(&d->d_dt, dims * sizeof(double)));
    DENGO_HIP_OK(hipMalloc(&d->d_ttot, dims * sizeof(double)));
    for (int **p : ip) DENGO_HIP_OK(hipMalloc(p, dims * sizeof(int)));
    d->cap = dims;
    return 0;
}

template <int MODE>
void kernel_for(bool persist, void (**k)(B::NetDev, double, double, const double *, double *,
                                         const double *, double *, long long, dengo_cell_outputs,
                                         int, unsigned long long *))
{
    if (persist) *k = B::k_big<MODE, true>;
    else *k = B::k_big<MODE, false>;
}

using KFn = void (*)(B::NetDev, double, double, const double *, double *, const double *,
                     double *, long long, dengo_cell_outputs, int, unsigned long long *);

// Enqueues the integration on stream; counter is the persistent schedule's
// work counter, events brackets the kernel with the struct's timing events.
int launch(synthetic_data *d, const dengo_options &o, double dtf, double dt0, const double *d_T,
           double *d_x, const double *d_rtol, double *d_atol, long long dims,
           const dengo_cell_outputs &out, hipStream_t stream, unsigned long long *counter,
           bool events)
{
    KFn k = nullptr;
    const bool persist = o.schedule == DENGO_PERSISTENT;
    const int mode = o.precision == DENGO_FP64 ? B::kLU : B::kGJ;
    if (mode == B::kLU) kernel_for<B::kLU>(persist, &k);
    else kernel_for<B::kGJ>(persist, &k);
    long long grid;
    if (persist) {
        grid = o.grid;
        if (grid <= 0) {
            int per_cu = 0;
            DENGO_HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, B::NTHR, 0));
            if (per_cu < 1) per_cu = 1;
            grid = (long long)per_cu * d->cu_count;
        }
        if (grid > dims) grid = dims;
    } else {
        grid = dims;
        if (grid > 0x7fffffffLL) return set_error("too many cells: %lld", dims);
    }
    if (events) DENGO_HIP_OK(hipEventRecord(d->ev0, stream));
    if (persist) {
        hipLaunchKernelGGL(k_reset_counter<0>, dim3(1), dim3(1), 0, stream, counter);
        DENGO_HIP_OK(hipGetLastError());
    }
    B::NetDev nv = net_view(d);
    nv.order = o.order;
    hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(B::NTHR), 0, stream, nv, dtf, dt0, d_T, d_x,
                       d_rtol, d_atol, dims, out, o.max_substeps, counter);
    DENGO_HIP_OK(hipGetLastError());
    if (events) DENGO_HIP_OK(hipEventRecord(d->ev1, stream));
    return 0;
}

int check_opts(const dengo_options *opt, dengo_options *o)
{
    dengo_default_options(o);
    if (opt) *o = *opt;
    if (o->block != 0 && o->block != B::NTHR)
        return set_error("block must be 0 or %d for the dense solver", B::NTHR);
    if (o->max_substeps <= 0) o->max_substeps = 50000;
    if (o->schedule != DENGO_PER_CELL && o->schedule != DENGO_PERSISTENT)
        return set_error("unknown schedule %d", o->schedule);
    if (o->precision != DENGO_FP64 && o->precision != DENGO_FP64_FAST)
        return set_error("precision %d is not available for the dense solver", o->precision);
    return 0;
}


int hook_prepare(void *data, const dengo_options *opt, int *full)
{
    synthetic_data *d = (synthetic_data *)data;
    dengo_options o;
    if (check_opts(opt, &o) || ensure_device(d)) return -1;
    KFn k = nullptr;
    const bool persist = true;
    if (o.precision == DENGO_FP64) kernel_for<B::kLU>(persist, &k);
    else kernel_for<B::kGJ>(persist, &k);
    int per_cu = 0;
    DENGO_HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, B::NTHR, 0));
    if (per_cu < 1) per_cu = 1;
    *full = per_cu * d->cu_count;
    return 0;
}

int hook_prepare_order(void *data, OrderBufs *ob, const long long *order, long long n,
                       void *stream)
{
    synthetic_data *d = (synthetic_data *)data;
    if (ensure_device(d)) return -1;
    if (!order || n <= 0) return 0;
    OrderBufs &bufs = ob ? *ob : d->ord;   // the dense kernel reads the order itself
    return order_alloc(bufs, n, 0) || order_check(bufs, order, n, (hipStream_t)stream) ? -1 : 0;
}

int hook_launch(void *data, const LaunchArgs &a)
{
    synthetic_data *d = (synthetic_data *)data;
    dengo_options o;
    if (check_opts(&a.opt, &o)) return -1;
    if (!a.T) return set_error("the dense network needs per-cell temperatures");
    return launch(d, o, a.dtf, a.dt0, a.T, a.x, a.rtol, a.atol, a.n, a.out, a.stream, a.counter,
                  false);
}
}  // namespace

extern "C" {

synthetic_data *synthetic_setup_data(int nspecies, int nreactions, const int *reactions,
                                     const double *A, const double *beta, const double *gamma)
{
    if (nspecies != B::N) {
        set_error("nspecies must be %d", B::N);
        return nullptr;
    }
    if (nreactions < 1 || nreactions > B::MAX_RXN) {
        set_error("nreactions must be in [1, %d]", B::MAX_RXN);
        return nullptr;
    }
    if (!reactions || !A || !beta || !gamma) {
        set_error("null argument");
        return nullptr;
    }
    for (int q = 0; q < 4 * nreactions; ++q)
        if (reactions[q] < 0 || reactions[q] >= nspecies) {
            set_error("reaction %d names species %d", q / 4, reactions[q]);
            return nullptr;
        }
    synthetic_data *d = new synthetic_data();
    d->n = nspecies;
    d->nrxn = nreactions;
    d->rxn.assign(reactions, reactions + 4 * nreactions);
    d->A.assign(A, A + nreactions);
    d->beta.assign(beta, beta + nreactions);
    d->gamma.assign(gamma, gamma + nreactions);
    d->prog = B::build_programs(reactions, nreactions);
    return d;
}

void synthetic_free_data(synthetic_data *d)
{
    if (!d) return;
    if (d->device >= 0) {
        void *ptrs[] = {d->d_rxn, d->d_A, d->d_beta, d->d_gamma, d->d_rhs, d->d_jac,
                        d->d_counter, d->d_red, d->d_x, d->d_T, d->d_rtol, d->d_atol,
                        d->d_dt, d->d_ttot, d->d_status, d->d_nsub, d->d_nrej, d->d_nsweep};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        if (d->ev0) (void)hipEventDestroy(d->ev0);
        if (d->ev1) (void)hipEventDestroy(d->ev1);
        order_free(d->ord);
    }
    delete d;
}

int dengo_synthetic_evolve_device(synthetic_data *d, const dengo_options *opt, double dtf,
                                  double dt0, const double *d_T, double *d_input,
                                  const double *d_rtol, double *d_atol, long long dims,
                                  const dengo_cell_outputs *outputs, double *ttot_min,
                                  double *dt_min, void *stream)
{
    if (!d || !d_T || !d_input || !d_rtol || !d_atol) return set_error("null argument");
    if (dims < 0) return set_error("negative cell count");
    dengo_options o;
    if (check_opts(opt, &o)) return -1;
    if (ensure_device(d)) return -1;
    dengo_cell_outputs out;
    memset(&out, 0, sizeof(out));
    if (outputs) out = *outputs;
    hipStream_t st = (hipStream_t)stream;
    if (ttot_min && (!out.ttot || !out.dt)) {
        if (ensure_scratch(d, dims)) return -1;
        if (!out.ttot) out.ttot = d->d_ttot;
        if (!out.dt) out.dt = d->d_dt;
    }
    if (o.order && dims > 0 &&
        (order_alloc(d->ord, dims, 0) || order_check(d->ord, o.order, dims, st)))
        return -1;
    if (dims > 0 && launch(d, o, dtf, dt0, d_T, d_input, d_rtol, d_atol, dims, out, st,
                           d->d_counter, true))
        return -1;
    if (ttot_min) {
        if (dims > 0) {
            hipLaunchKernelGGL(k_min_ttot<0>, dim3(1), dim3(1024), 0, st, out.ttot, out.dt, dims,
                               d->d_red);
            DENGO_HIP_OK(hipGetLastError());
            double red[2];
            DENGO_HIP_OK(hipMemcpyAsync(red, d->d_red, sizeof(red), hipMemcpyDeviceToHost, st));
            DENGO_HIP_OK(hipStreamSynchronize(st));
            *ttot_min = red[0];
            if (dt_min) *dt_min = red[1];
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, d->ev0, d->ev1) == hipSuccess) set_last_ms(ms);
        } else {
            *ttot_min = 0.0;
            if (dt_min) *dt_min = dt0;
        }
    }
    return 0;
}

double dengo_synthetic_evolve_host(synthetic_data *d, const dengo_options *opt, double dtf,
                                   double *dt, const double *T, double *input, double *rtol,
                                   double *atol, long long dims, int *status, int *nsub,
                                   int *nrej, int *nsweep, double *dt_cell)
{
    if (!d || !dt || !T || !input || !rtol || !atol) {
        set_error("null argument");
        return -1.0;
    }
    if (dims <= 0) return 0.0;
    if (ensure_device(d) || ensure_scratch(d, dims)) return -1.0;
    const size_t bytes = (size_t)dims * B::N * sizeof(double);
    if (hipMemcpy(d->d_x, input, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d->d_rtol, rtol, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d->d_atol, atol, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d->d_T, T, dims * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        set_error("host-to-device copy failed");
        return -1.0;
    }
    dengo_cell_outputs out;
    memset(&out, 0, sizeof(out));
    out.status = d->d_status;
    out.nsub = d->d_nsub;
    out.nrej = d->d_nrej;
    out.nsweep = d->d_nsweep;
    out.dt = d->d_dt;
    out.ttot = d->d_ttot;
    double tmin = 0.0, dtm = *dt;
    if (dengo_synthetic_evolve_device(d, opt, dtf, *dt, d->d_T, d->d_x, d->d_rtol, d->d_atol, dims,
                                      &out, &tmin, &dtm, NULL) != 0)
        return -1.0;
    struct Copy { void *dst; const void *src; size_t n; };
    const Copy cps[] = {{input, d->d_x, bytes}, {atol, d->d_atol, bytes},
                        {status, d->d_status, dims * sizeof(int)},
                        {nsub, d->d_nsub, dims * sizeof(int)},
                        {nrej, d->d_nrej, dims * sizeof(int)},
                        {nsweep, d->d_nsweep, dims * sizeof(int)},
                        {dt_cell, d->d_dt, dims * sizeof(double)}};
    for (const Copy &c : cps) {
        if (c.dst && hipMemcpy(c.dst, c.src, c.n, hipMemcpyDeviceToHost) != hipSuccess) {
            set_error("device-to-host copy failed");
            return -1.0;
        }
    }
    *dt = dtm;
    return tmin;
}

double dengo_evolve_synthetic(double dtf, double *dt, const double *T, double *input,
                              double *rtol, double *atol, long long dims, synthetic_data *data)
{
    return dengo_synthetic_evolve_host(data, nullptr, dtf, dt, T, input, rtol, atol, dims, nullptr,
                                       nullptr, nullptr, nullptr, nullptr);
}

long long dengo_synthetic_eval(synthetic_data *d, const double *T, const double *u,
                               long long dims, double *f, double *J)
{
    if (!d || !T || !u || !f) return set_error("null argument");
    if (dims <= 0) return 0;
    if (ensure_device(d)) return -1;
    const size_t bu = (size_t)dims * B::N * sizeof(double);
    double *du = nullptr, *df = nullptr, *dJ = nullptr, *dT = nullptr;
    int *dn = nullptr;
    DENGO_HIP_OK(hipMalloc(&du, bu));
    DENGO_HIP_OK(hipMalloc(&df, bu));
    DENGO_HIP_OK(hipMalloc(&dT, dims * sizeof(double)));
    DENGO_HIP_OK(hipMalloc(&dn, dims * sizeof(int)));
    if (J) DENGO_HIP_OK(hipMalloc(&dJ, bu * B::N));
    DENGO_HIP_OK(hipMemcpy(du, u, bu, hipMemcpyHostToDevice));
    DENGO_HIP_OK(hipMemcpy(dT, T, dims * sizeof(double), hipMemcpyHostToDevice));
    DENGO_HIP_OK(hipMemcpy(df, f, bu, hipMemcpyHostToDevice));
    if (J) DENGO_HIP_OK(hipMemcpy(dJ, J, bu * B::N, hipMemcpyHostToDevice));
    if (dims > 0x7fffffffLL) return set_error("too many cells: %lld", dims);
    hipLaunchKernelGGL(B::k_big_eval, dim3((unsigned)dims), dim3(B::NTHR), 0, 0, net_view(d), dT, du,
                       dims, df, dJ, dn);
    DENGO_HIP_OK(hipGetLastError());
    std::vector<int> neg(dims);
    DENGO_HIP_OK(hipMemcpy(neg.data(), dn, dims * sizeof(int), hipMemcpyDeviceToHost));
    DENGO_HIP_OK(hipMemcpy(f, df, bu, hipMemcpyDeviceToHost));
    if (J) DENGO_HIP_OK(hipMemcpy(J, dJ, bu * B::N, hipMemcpyDeviceToHost));
    void *ptrs[] = {du, df, dJ, dT, dn};
    for (void *q : ptrs)
        if (q) DENGO_HIP_OK(hipFree(q));
    long long nneg = 0;
    for (int v : neg) nneg += v;
    return nneg;
}

}  // extern "C"

NetHooks dengo::hooks_synthetic()
{
    NetHooks h;
    h.nspecies = B::N;
    h.isothermal = 1;
    h.prepare = hook_prepare;
    h.prepare_order = hook_prepare_order;
    h.resident = nullptr;
    h.launch = hook_launch;
    return h;
}
