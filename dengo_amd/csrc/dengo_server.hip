// Patch server: continuous batching of chemistry solves on one GPU.
//
// Host-side runtime around the per-network kernels.  One worker thread owns
// every HIP call of the server:
//   reap    a slot whose kernel has finished gets its copy-back enqueued; a
//           slot whose copy-back has finished completes its request, returns
//           its arena block and becomes free;
//   admit   while a slot is free, the first waiting request whose buffers fit
//           in the arena (pick_admission) is copied in and launched on that
//           slot's stream, with the persistent grid capped at
//           full_grid / grid_share so that several patches run at once;
//   idle    with nothing in flight and nothing waiting the worker sleeps on a
//           condition variable; otherwise it polls its events every ~20 us.
// Clients only queue requests (submit) and collect completion records (poll).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "arena.h"
#include "dengo_server.h"
#include "dengo_solver.h"

using namespace dengo;

namespace {

// The three built-in networks take ids 0-2 (enum dengo_network); networks
// added with dengo_server_add_network take the next ids.
constexpr int kMaxNets = 16;
constexpr int kBuiltinNets = 3;

struct Request {
    long long ticket = -1;
    int net = 0;
    bool device = false;   // the arrays are the caller's device memory: no copies
    long long n = 0;
    double *x = nullptr;
    const double *rtol = nullptr;
    double *atol = nullptr;
    const double *T = nullptr;
    int *status = nullptr;
    double dtf = 0, dt0 = -1, z = -1;
    double t_submit = 0;
    size_t bytes = 0;
    int bypassed = 0;
    std::vector<long long> order;   // dense network, host submits: densest first
    size_t sort_temp = 0;           // dense network, device submits: the sort's scratch
};

// Densest cells first: indices by descending total of the n x ns species
// columns, a NaN total counting as -inf, ties in index order
// (runtime.density_order follows the same rule).
std::vector<long long> density_order(const double *x, long long n, int ns)
{
    std::vector<double> key((size_t)n);
    std::vector<long long> ord((size_t)n);
    for (long long i = 0; i < n; ++i) {
        double t = 0.0;
        for (int j = 0; j < ns; ++j) t += x[(size_t)i * ns + j];
        key[(size_t)i] = std::isnan(t) ? -HUGE_VAL : t;
        ord[(size_t)i] = i;
    }
    std::stable_sort(ord.begin(), ord.end(),
                     [&key](long long a, long long b) { return key[(size_t)a] > key[(size_t)b]; });
    return ord;
}

enum SlotPhase { kFree = 0, kRunning = 1, kCopyBack = 2 };

struct Slot {
    hipStream_t stream = nullptr;
    unsigned long long *counter = nullptr;
    hipEvent_t ev_start = nullptr, ev_end = nullptr, ev_done = nullptr;
    int phase = kFree;
    Request req;
    long long off = -1;
    double *dx = nullptr, *drt = nullptr, *dat = nullptr, *dT = nullptr;
    int *dst = nullptr;
    long long *dord = nullptr;         // dense network: the hand-out order
    std::vector<long long> hord;       // its host copy, alive until the slot completes
    int *dcounts = nullptr;            // device submits: status counts, device and
    int *hcounts = nullptr;            // pinned host copies
    double t_admit = 0;
    int inflight_at_admit = 0, max_inflight = 0;
};

size_t round_up(size_t n) { return Arena::round_up(n); }

// Cells of a device-submitted patch by final status (0, 1, 2); one workgroup.
template <int>
__global__ void __launch_bounds__(256) k_status_counts(const int *status, long long n, int *out)
{
    __shared__ int c[3][256];
    int a = 0, b = 0, d = 0;
    for (long long i = threadIdx.x; i < n; i += 256) {
        const int v = status[i];
        a += v == 0;
        b += v == 1;
        d += v == 2;
    }
    c[0][threadIdx.x] = a;
    c[1][threadIdx.x] = b;
    c[2][threadIdx.x] = d;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; ++k) c[k][threadIdx.x] += c[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 3) out[threadIdx.x] = c[threadIdx.x][0];
}

// Sort keys of a device-resident dense patch: each cell's total density (the
// species columns summed in index order, as density_order does on the host; a
// NaN total counts as -inf) and its index.
template <int>
__global__ void __launch_bounds__(256) k_density_keys(const double *x, long long n, int ns,
                                                      double *key, long long *idx)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double t = 0.0;
    for (int j = 0; j < ns; ++j) t += x[i * ns + j];
    key[i] = isnan(t) ? -HUGE_VAL : t;
    idx[i] = i;
}

// Scratch bytes of the stable descending radix sort of n (key, index) pairs.
size_t sort_temp_bytes(long long n)
{
    size_t b = 0;
    if (hipcub::DeviceRadixSort::SortPairsDescending(nullptr, b, (const double *)nullptr,
                                                     (double *)nullptr,
                                                     (const long long *)nullptr,
                                                     (long long *)nullptr, (int)n) != hipSuccess)
        return 0;
    return b;
}

// A device pointer the caller hands over must be device memory of at least
// `bytes` from ptr on; anything else would fault the kernel.
int check_device_range(const void *ptr, size_t bytes, const char *what)
{
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, ptr) != hipSuccess) {
        (void)hipGetLastError();
        return set_error("%s is not device memory", what);
    }
    if (at.type != hipMemoryTypeDevice)
        return set_error("%s is not device memory", what);
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)ptr) != hipSuccess) {
        (void)hipGetLastError();
        return set_error("%s: its allocation is unknown", what);
    }
    if ((const char *)ptr + bytes > (const char *)base + size)
        return set_error("%s holds fewer than %zu bytes", what, bytes);
    return 0;
}

}  // namespace

struct dengo_server {
    int device = 0;
    dengo_server_config cfg;
    int precision = DENGO_FP64;
    void *data[kMaxNets] = {};
    NetHooks hooks[kMaxNets];
    int grid_cap[kMaxNets] = {};
    int prec[kMaxNets];                // per network: precision, or -1 for `precision`
    int nnets = kBuiltinNets;
    char *arena_base = nullptr;
    Arena arena;
    std::vector<Slot> slots;
    unsigned long long *counters = nullptr;
    hipEvent_t epoch = nullptr;

    std::mutex mu;
    std::mutex add_mu;                 // serialises dengo_server_add_network
    std::condition_variable cv_work, cv_done;
    std::deque<Request> waiting;
    std::deque<dengo_patch_result> completed;
    long long next_ticket = 0, submitted = 0, ncompleted = 0, nfailed = 0, cells_completed = 0;
    int inflight = 0, max_inflight_seen = 0;
    double first_submit = -1, last_done = -1;
    bool stop = false;
    std::thread worker;
    std::chrono::steady_clock::time_point t0;

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    int net_precision(int net) const
    {
        const int pr = prec[net] >= 0 ? prec[net] : precision;
        if (hooks[net].isothermal && pr == DENGO_MIXED) return DENGO_FP64_FAST;
        return pr;
    }

    size_t request_bytes(int net, long long n, bool device, size_t sort_temp) const
    {
        // a device patch is integrated in the caller's buffers; a dense one
        // still needs its densest-first order: keys in and out, indices in,
        // the order, and the sort's scratch
        if (device)
            return hooks[net].isothermal && cfg.dense_order
                       ? 4 * round_up((size_t)n * sizeof(double)) + round_up(sort_temp)
                       : 0;
        const size_t row = (size_t)n * hooks[net].nspecies * sizeof(double);
        size_t b = 3 * round_up(row) + round_up((size_t)n * sizeof(int));
        // the dense network's patches also carry T[n] and the order[n]
        if (hooks[net].isothermal) b += 2 * round_up((size_t)n * sizeof(double));
        return b;
    }

    // Makes network net's tables resident and caps its persistent grid at
    // 1/share of the grid that fills the GPU (share 0: cfg.grid_share).
    int prepare_net(int net, int share)
    {
        dengo_options o;
        dengo_default_options(&o);
        o.precision = net_precision(net);
        o.block = 0;
        int full = 0;
        if (hooks[net].prepare(data[net], &o, &full)) return -1;
        if (share <= 0) share = cfg.grid_share;
        grid_cap[net] = full / share > 0 ? full / share : 1;
        return 0;
    }

    // Called with mu held; the slot's device work has finished (or failed).
    void complete(Slot &s, int err, const char *msg)
    {
        dengo_patch_result r;
        memset(&r, 0, sizeof(r));
        r.ticket = s.req.ticket;
        r.network = s.req.net;
        r.error = err;
        r.ncells = s.req.n;
        r.slot = (int)(&s - slots.data());
        r.inflight_at_admit = s.inflight_at_admit;
        r.max_inflight = s.max_inflight;
        r.t_submit = s.req.t_submit;
        r.t_admit = s.t_admit;
        r.t_done = now();
        r.latency_ms = 1e3 * (r.t_done - r.t_submit);
        r.queue_ms = 1e3 * (r.t_admit - r.t_submit);
        if (err == 0) {
            float a = 0.f, b = 0.f;
            if (hipEventElapsedTime(&a, epoch, s.ev_start) == hipSuccess &&
                hipEventElapsedTime(&b, epoch, s.ev_end) == hipSuccess) {
                r.dev_start_ms = a;
                r.dev_end_ms = b;
            }
            if (s.req.device) {
                if (s.req.status)
                    for (int v = 0; v < 3; ++v) r.status_counts[v] = s.hcounts[v];
            } else if (s.req.status) {
                for (long long i = 0; i < s.req.n; ++i) {
                    const int v = s.req.status[i];
                    if (v >= 0 && v < 3) r.status_counts[v]++;
                }
            }
            ncompleted++;
            cells_completed += s.req.n;
        } else {
            nfailed++;
            snprintf(r.message, sizeof(r.message), "%s", msg ? msg : "failed");
        }
        last_done = r.t_done;
        completed.push_back(r);
        if (s.off >= 0) arena.release((size_t)s.off, s.req.bytes);
        s.off = -1;
        s.phase = kFree;
        inflight--;
        cv_done.notify_all();
    }

    // Enqueues copy-in, kernel and the start/end events; mu not held.
    int enqueue_solve(Slot &s)
    {
        const Request &q = s.req;
        const size_t row = (size_t)q.n * hooks[q.net].nspecies * sizeof(double);
        if (q.device) {
            if (s.dord) {
                // densest cells first, sorted on the slot's stream (stable:
                // equal totals keep index order, as on the host)
                const size_t nb = round_up((size_t)q.n * sizeof(double));
                char *p = arena_base + s.off;
                double *kin = (double *)p, *kout = (double *)(p + nb);
                long long *iin = (long long *)(p + 2 * nb);
                void *tmp = p + 4 * nb;
                const unsigned blocks = (unsigned)((q.n + 255) / 256);
                hipLaunchKernelGGL(k_density_keys<0>, dim3(blocks), dim3(256), 0, s.stream, q.x,
                                   q.n, hooks[q.net].nspecies, kin, iin);
                DENGO_HIP_OK(hipGetLastError());
                size_t tb = q.sort_temp;
                DENGO_HIP_OK(hipcub::DeviceRadixSort::SortPairsDescending(
                    tmp, tb, kin, kout, iin, s.dord, (int)q.n, 0, 64, s.stream));
            }
            return enqueue_kernel(s);
        }
        DENGO_HIP_OK(hipMemcpyAsync(s.dx, q.x, row, hipMemcpyHostToDevice, s.stream));
        DENGO_HIP_OK(hipMemcpyAsync(s.drt, q.rtol, row, hipMemcpyHostToDevice, s.stream));
        DENGO_HIP_OK(hipMemcpyAsync(s.dat, q.atol, row, hipMemcpyHostToDevice, s.stream));
        if (s.dT)
            DENGO_HIP_OK(hipMemcpyAsync(s.dT, q.T, (size_t)q.n * sizeof(double),
                                        hipMemcpyHostToDevice, s.stream));
        if (s.dord)   // the order was built by the submitting thread
            DENGO_HIP_OK(hipMemcpyAsync(s.dord, s.hord.data(), (size_t)q.n * sizeof(long long),
                                        hipMemcpyHostToDevice, s.stream));
        return enqueue_kernel(s);
    }

    // The integration itself, on the slot's buffers, bracketed by the
    // slot's start/end events.
    int enqueue_kernel(Slot &s)
    {
        const Request &q = s.req;
        DENGO_HIP_OK(hipEventRecord(s.ev_start, s.stream));
        LaunchArgs a;
        memset(&a, 0, sizeof(a));
        dengo_default_options(&a.opt);
        a.opt.precision = net_precision(q.net);
        a.opt.schedule = DENGO_PERSISTENT;
        a.opt.block = 0;
        a.opt.grid = grid_cap[q.net];
        a.dtf = q.dtf;
        a.dt0 = q.dt0;
        a.z = q.z;
        a.T = s.dT;
        a.x = s.dx;
        a.rtol = s.drt;
        a.atol = s.dat;
        a.n = q.n;
        a.out.status = s.dst;
        a.stream = s.stream;
        a.counter = s.counter;
        a.opt.order = s.dord;
        if (hooks[q.net].launch(data[q.net], a)) return -1;
        DENGO_HIP_OK(hipEventRecord(s.ev_end, s.stream));
        return 0;
    }

    int enqueue_copy_back(Slot &s)
    {
        const Request &q = s.req;
        if (q.device) {
            // the state stays where it is; only the status counts come back
            if (q.status) {
                hipLaunchKernelGGL(k_status_counts<0>, dim3(1), dim3(256), 0, s.stream, q.status,
                                   q.n, s.dcounts);
                DENGO_HIP_OK(hipGetLastError());
                DENGO_HIP_OK(hipMemcpyAsync(s.hcounts, s.dcounts, 3 * sizeof(int),
                                            hipMemcpyDeviceToHost, s.stream));
            }
            DENGO_HIP_OK(hipEventRecord(s.ev_done, s.stream));
            return 0;
        }
        const size_t row = (size_t)q.n * hooks[q.net].nspecies * sizeof(double);
        DENGO_HIP_OK(hipMemcpyAsync(q.x, s.dx, row, hipMemcpyDeviceToHost, s.stream));
        DENGO_HIP_OK(hipMemcpyAsync(q.atol, s.dat, row, hipMemcpyDeviceToHost, s.stream));
        if (q.status)
            DENGO_HIP_OK(hipMemcpyAsync(q.status, s.dst, (size_t)q.n * sizeof(int),
                                        hipMemcpyDeviceToHost, s.stream));
        DENGO_HIP_OK(hipEventRecord(s.ev_done, s.stream));
        return 0;
    }

    void fail_slot(std::unique_lock<std::mutex> &lk, Slot &s)
    {
        char msg[160];
        snprintf(msg, sizeof(msg), "%s", last_error_text());
        lk.unlock();
        (void)hipStreamSynchronize(s.stream);   // nothing may still touch the block
        lk.lock();
        complete(s, -1, msg);
    }

    void run()
    {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            if (stop && waiting.empty() && inflight == 0) break;
            // reap
            for (Slot &s : slots) {
                if (s.phase == kRunning) {
                    const hipError_t e = hipEventQuery(s.ev_end);
                    if (e == hipErrorNotReady) continue;
                    if (e != hipSuccess) {
                        set_error("kernel of ticket %lld: %s", s.req.ticket, hipGetErrorString(e));
                        fail_slot(lk, s);
                        continue;
                    }
                    s.phase = kCopyBack;
                    lk.unlock();
                    const int rc = enqueue_copy_back(s);
                    lk.lock();
                    if (rc) fail_slot(lk, s);
                } else if (s.phase == kCopyBack) {
                    const hipError_t e = hipEventQuery(s.ev_done);
                    if (e == hipErrorNotReady) continue;
                    if (e != hipSuccess) {
                        set_error("copy-back of ticket %lld: %s", s.req.ticket,
                                  hipGetErrorString(e));
                        fail_slot(lk, s);
                        continue;
                    }
                    complete(s, 0, nullptr);
                }
            }
            // admit
            while (!waiting.empty()) {
                Slot *fs = nullptr;
                for (Slot &s : slots)
                    if (s.phase == kFree) {
                        fs = &s;
                        break;
                    }
                if (!fs) break;
                const int idx = pick_admission(
                    waiting, [&](const Request &r) { return r.bytes == 0 || arena.fits(r.bytes); },
                    cfg.max_bypass);
                if (idx < 0) break;
                Slot &s = *fs;
                s.req = waiting[idx];
                waiting.erase(waiting.begin() + idx);
                if (s.req.device) {
                    // the caller's buffers; the dense network's cells keep
                    // index order (the densest-first order is built from host
                    // state)
                    s.off = s.req.bytes ? arena.alloc(s.req.bytes) : -1;
                    s.dx = s.req.x;
                    s.drt = const_cast<double *>(s.req.rtol);
                    s.dat = s.req.atol;
                    s.dst = s.req.status;
                    s.dT = const_cast<double *>(s.req.T);
                    s.dord = s.req.bytes ? (long long *)(arena_base + s.off +
                                                         3 * round_up((size_t)s.req.n * 8))
                                         : nullptr;
                } else {
                    s.off = arena.alloc(s.req.bytes);
                    char *p = arena_base + s.off;
                    const size_t row = round_up((size_t)s.req.n * hooks[s.req.net].nspecies * 8);
                    s.dx = (double *)p;
                    s.drt = (double *)(p + row);
                    s.dat = (double *)(p + 2 * row);
                    s.dst = (int *)(p + 3 * row);
                    const size_t st_end = 3 * row + round_up((size_t)s.req.n * sizeof(int));
                    s.dT = hooks[s.req.net].isothermal ? (double *)(p + st_end) : nullptr;
                    s.dord = hooks[s.req.net].isothermal && cfg.dense_order
                                 ? (long long *)(p + st_end +
                                                 round_up((size_t)s.req.n * sizeof(double)))
                                 : nullptr;
                    s.hord = std::move(s.req.order);
                }
                s.phase = kRunning;
                s.t_admit = now();
                inflight++;
                if (inflight > max_inflight_seen) max_inflight_seen = inflight;
                s.inflight_at_admit = inflight;
                for (Slot &o : slots)
                    if (o.phase != kFree && o.max_inflight < inflight) o.max_inflight = inflight;
                lk.unlock();
                const int rc = enqueue_solve(s);
                lk.lock();
                if (rc) fail_slot(lk, s);
            }
            if (inflight == 0 && waiting.empty()) {
                if (stop) break;
                cv_work.wait(lk, [this] { return stop || !waiting.empty(); });
            } else {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
                lk.lock();
            }
        }
    }
};

extern "C" {

void dengo_server_default_config(dengo_server_config *cfg)
{
    memset(cfg, 0, sizeof(*cfg));
    cfg->hbm_budget = (size_t)8 << 30;
    cfg->slots = 4;
    cfg->grid_share = 0;
    cfg->max_bypass = 8;
    for (int k = 0; k < 3; ++k) cfg->precision[k] = -1;
    for (int k = 0; k < 3; ++k) cfg->grid_shares[k] = 0;
    cfg->dense_order = 1;
}

dengo_server *dengo_server_create(const dengo_server_config *cfg_in, void *primordial,
                                  void *oxygen, void *synthetic, int precision)
{
    dengo_server_config cfg;
    dengo_server_default_config(&cfg);
    if (cfg_in) cfg = *cfg_in;
    if (cfg.slots < 1 || cfg.slots > 64) {
        set_error("slots must be in [1, 64]");
        return nullptr;
    }
    if (cfg.grid_share <= 0) cfg.grid_share = cfg.slots;
    if (cfg.max_bypass <= 0) cfg.max_bypass = 8;
    if (cfg.hbm_budget < Arena::kAlign) {
        set_error("hbm_budget too small");
        return nullptr;
    }
    if (precision != DENGO_FP64 && precision != DENGO_MIXED && precision != DENGO_FP64_FAST) {
        set_error("unknown precision %d", precision);
        return nullptr;
    }
    for (int k = 0; k < kBuiltinNets; ++k)
        if (cfg.grid_shares[k] < 0 || cfg.grid_shares[k] > 1024) {
            set_error("grid_shares[%d] must be in [0, 1024]", k);
            return nullptr;
        }
    for (int k = 0; k < kBuiltinNets; ++k)
        if (cfg.precision[k] > DENGO_FP64_FAST || cfg.precision[k] < -1) {
            set_error("unknown precision %d for network %d", cfg.precision[k], k);
            return nullptr;
        }
    dengo_server *s = new dengo_server();
    s->cfg = cfg;
    s->precision = precision;
    for (int k = 0; k < kMaxNets; ++k) s->prec[k] = k < kBuiltinNets ? cfg.precision[k] : -1;
    s->t0 = std::chrono::steady_clock::now();
    s->data[DENGO_NET_PRIMORDIAL] = primordial;
    s->data[DENGO_NET_OXYGEN] = oxygen;
    s->data[DENGO_NET_SYNTHETIC] = synthetic;
    s->hooks[DENGO_NET_PRIMORDIAL] = hooks_primordial();
    s->hooks[DENGO_NET_OXYGEN] = hooks_oxygen();
    s->hooks[DENGO_NET_SYNTHETIC] = hooks_synthetic();
    auto fail = [&](const char *what, hipError_t e) -> dengo_server * {
        if (e != hipSuccess) set_error("%s: %s", what, hipGetErrorString(e));
        dengo_server_destroy(s);
        return nullptr;
    };
    hipError_t e = hipGetDevice(&s->device);
    if (e != hipSuccess) return fail("hipGetDevice", e);
    for (int k = 0; k < kBuiltinNets; ++k) {
        if (!s->data[k]) continue;
        if (s->prepare_net(k, cfg.grid_shares[k])) return fail("prepare", hipSuccess);
    }
    if ((e = hipMalloc(&s->arena_base, cfg.hbm_budget)) != hipSuccess)
        return fail("arena hipMalloc", e);
    s->arena.reset(cfg.hbm_budget);
    if ((e = hipMalloc(&s->counters, 64 * sizeof(unsigned long long) * cfg.slots)) != hipSuccess)
        return fail("counter hipMalloc", e);
    if ((e = hipEventCreate(&s->epoch)) != hipSuccess) return fail("hipEventCreate", e);
    s->slots.resize(cfg.slots);
    for (int i = 0; i < cfg.slots; ++i) {
        Slot &sl = s->slots[i];
        sl.counter = s->counters + 64 * i;   // one 512-byte line per slot
        if ((e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking)) != hipSuccess ||
            (e = hipEventCreate(&sl.ev_start)) != hipSuccess ||
            (e = hipEventCreate(&sl.ev_end)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&sl.ev_done, hipEventDisableTiming)) != hipSuccess ||
            (e = hipMalloc(&sl.dcounts, 4 * sizeof(int))) != hipSuccess ||
            (e = hipHostMalloc(&sl.hcounts, 4 * sizeof(int), hipHostMallocDefault)) != hipSuccess)
            return fail("slot set-up", e);
    }
    if ((e = hipEventRecord(s->epoch, s->slots[0].stream)) != hipSuccess ||
        (e = hipEventSynchronize(s->epoch)) != hipSuccess)
        return fail("epoch event", e);
    s->worker = std::thread([s] { s->run(); });
    return s;
}

int dengo_server_add_network(dengo_server *s, void *data, const dengo_net_hooks *hooks,
                             int precision, int grid_share)
{
    if (!s || !data || !hooks) return set_error("null argument");
    if (precision < -1 || precision > DENGO_FP64_FAST)
        return set_error("unknown precision %d", precision);
    if (grid_share < 0 || grid_share > 1024) return set_error("grid_share must be in [0, 1024]");
    const NetHooks &h = *reinterpret_cast<const NetHooks *>(hooks);
    if (h.nspecies < 1 || h.nspecies > 1024 || !h.prepare || !h.launch)
        return set_error("not a network's server hooks");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != s->device)
        return set_error("add the network on the server's device (%d)", s->device);
    // one registration at a time; the worker reads only registered entries
    // (ids below nnets), so the new one is filled in without holding mu
    std::lock_guard<std::mutex> ga(s->add_mu);
    int id;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) return set_error("server is shutting down");
        if (s->nnets >= kMaxNets) return set_error("a server holds at most %d networks", kMaxNets);
        id = s->nnets;
    }
    s->data[id] = data;
    s->hooks[id] = h;
    s->prec[id] = precision;
    if (s->prepare_net(id, grid_share)) {
        s->data[id] = nullptr;
        return set_error("network %d: its tables could not be made resident", id);
    }
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->nnets = id + 1;
    }
    return id;
}

static long long submit_request(dengo_server *s, int network, long long ncells, double *state,
                                const double *rtol, double *atol, const double *T, int *status,
                                double dtf, double dt0, double z, bool device)
{
    if (!s) return set_error("null server");
    if (ncells <= 0) return set_error("a patch needs at least one cell");
    if (!state || !rtol || !atol) return set_error("null argument");
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (network < 0 || network >= s->nnets || !s->data[network])
            return set_error("network %d is not served", network);
    }
    const NetHooks &h = s->hooks[network];
    if (h.isothermal && !T) return set_error("the synthetic network needs per-cell temperatures");
    if (device) {
        const size_t row = (size_t)ncells * h.nspecies * sizeof(double);
        if (check_device_range(state, row, "state") || check_device_range(rtol, row, "rtol") ||
            check_device_range(atol, row, "atol") ||
            (h.isothermal && check_device_range(T, (size_t)ncells * sizeof(double), "T")) ||
            (status && check_device_range(status, (size_t)ncells * sizeof(int), "status")))
            return -1;
    }
    Request r;
    r.device = device;
    r.net = network;
    r.n = ncells;
    r.x = state;
    r.rtol = rtol;
    r.atol = atol;
    r.T = s->hooks[network].isothermal ? T : nullptr;
    r.status = status;
    r.dtf = dtf;
    r.dt0 = dt0;
    r.z = z;
    if (device && h.isothermal && s->cfg.dense_order) {
        if (ncells > 0x7fffffffLL) return set_error("a device patch holds at most 2^31 - 1 cells");
        r.sort_temp = sort_temp_bytes(ncells);
        if (r.sort_temp == 0) return set_error("cannot size the cell sort of %lld cells", ncells);
    }
    r.bytes = s->request_bytes(network, ncells, device, r.sort_temp);
    // the densest-first order of a dense host patch is built here, on the
    // caller's thread, so that the worker never stalls on it
    if (!device && h.isothermal && s->cfg.dense_order)
        r.order = density_order(state, ncells, h.nspecies);
    if (r.bytes > s->arena.size())
        return set_error("a patch of %lld cells needs %zu bytes; the arena holds %zu", ncells,
                         r.bytes, s->arena.size());
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) return set_error("server is shutting down");
        r.ticket = s->next_ticket++;
        r.t_submit = s->now();
        if (s->first_submit < 0) s->first_submit = r.t_submit;
        s->submitted++;
        s->waiting.push_back(r);
    }
    s->cv_work.notify_one();
    return r.ticket;
}

long long dengo_server_submit(dengo_server *s, int network, long long ncells, double *state,
                              const double *rtol, double *atol, const double *T, int *status,
                              double dtf, double dt0, double z)
{
    return submit_request(s, network, ncells, state, rtol, atol, T, status, dtf, dt0, z, false);
}

long long dengo_server_submit_device(dengo_server *s, int network, long long ncells,
                                     double *d_state, const double *d_rtol, double *d_atol,
                                     const double *d_T, int *d_status, double dtf, double dt0,
                                     double z)
{
    return submit_request(s, network, ncells, d_state, d_rtol, d_atol, d_T, d_status, dtf, dt0, z,
                          true);
}

int dengo_server_poll(dengo_server *s, dengo_patch_result *out, int max)
{
    if (!s || !out || max <= 0) return 0;
    std::lock_guard<std::mutex> g(s->mu);
    int k = 0;
    while (k < max && !s->completed.empty()) {
        out[k++] = s->completed.front();
        s->completed.pop_front();
    }
    return k;
}

int dengo_server_drain(dengo_server *s, double timeout_s)
{
    if (!s) return set_error("null server");
    std::unique_lock<std::mutex> lk(s->mu);
    const auto until = std::chrono::steady_clock::now() +
                       std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                           std::chrono::duration<double>(timeout_s));
    const bool ok = s->cv_done.wait_until(lk, until, [s] {
        return s->ncompleted + s->nfailed == s->submitted;
    });
    if (!ok)
        return set_error("drain timed out: %lld of %lld requests done",
                         s->ncompleted + s->nfailed, s->submitted);
    return 0;
}

void dengo_server_get_stats(dengo_server *s, dengo_server_stats *st)
{
    if (!s || !st) return;
    std::lock_guard<std::mutex> g(s->mu);
    memset(st, 0, sizeof(*st));
    st->submitted = s->submitted;
    st->completed = s->ncompleted;
    st->failed = s->nfailed;
    st->waiting = (long long)s->waiting.size();
    st->inflight = s->inflight;
    st->cells_completed = s->cells_completed;
    st->arena_bytes = s->arena.size();
    st->arena_used = s->arena.used();
    st->arena_peak = s->arena.peak();
    st->slots = (int)s->slots.size();
    st->max_inflight_seen = s->max_inflight_seen;
    st->first_submit = s->first_submit;
    st->last_done = s->last_done;
}

double dengo_server_now(dengo_server *s) { return s ? s->now() : 0.0; }

void dengo_server_destroy(dengo_server *s)
{
    if (!s) return;
    if (s->worker.joinable()) {
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        s->worker.join();
    }
    for (Slot &sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.ev_start) (void)hipEventDestroy(sl.ev_start);
        if (sl.ev_end) (void)hipEventDestroy(sl.ev_end);
        if (sl.ev_done) (void)hipEventDestroy(sl.ev_done);
        if (sl.dcounts) (void)hipFree(sl.dcounts);
        if (sl.hcounts) (void)hipHostFree(sl.hcounts);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    if (s->epoch) (void)hipEventDestroy(s->epoch);
    if (s->counters) (void)hipFree(s->counters);
    if (s->arena_base) (void)hipFree(s->arena_base);
    delete s;
}

void *dengo_host_alloc(size_t bytes)
{
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        set_error("hipHostMalloc of %zu bytes failed", bytes);
        return nullptr;
    }
    return p;
}

void dengo_host_free(void *p)
{
    if (p) (void)hipHostFree(p);
}

}  // extern "C"
