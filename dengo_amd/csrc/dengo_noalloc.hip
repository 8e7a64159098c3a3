// dengo_primordial_evolve_device_noalloc (include/dengo_noalloc.h): the
// device-pointer integration of dengo_primordial_evolve_device for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops).  The kernels and their launch are the same; the only
// scratch they need -- the persistent schedule's work counter, the per-cell
// elapsed times and final substeps that the smallest-elapsed-time reduction
// reads when the caller gives no such outputs, and the reduction's result --
// lives in the file-scope array `scratch` below.  Each GPU holds its own copy
// of that array from module load until the process ends; its address on the
// current device comes from hipGetSymbolAddress, looked up once per device on
// the first call there.  Nothing here calls hipMalloc or hipFree.
#include <mutex>

#include "dengo_noalloc.h"
#include "dengo_solver.h"

using namespace dengo;

namespace {

// Largest cell count whose per-cell elapsed times and final substeps fit the
// scratch array: 128^3 cells.
constexpr long long kMaxCells = DENGO_NOALLOC_MAX_CELLS;
// Layout: work counter at byte 0, reduction result (ttot, dt) at byte 64,
// per-cell ttot from byte kHeader, per-cell dt after it.
constexpr size_t kHeader = 256;
constexpr size_t kBytes = kHeader + 2 * (size_t)kMaxCells * sizeof(double);

static __device__ __align__(256) unsigned char scratch[kBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char *g_scratch[kMaxDevices];   // per device: the address of `scratch` there

// The address of `scratch` on the current device (hipGetSymbolAddress with
// that device current, once per device).
int device_scratch(unsigned char **out)
{
    int dev = -1;
    DENGO_HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return set_error("device %d is beyond the %d devices this entry point tracks", dev,
                         kMaxDevices);
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                             hipGetErrorString(e));
        g_scratch[dev] = (unsigned char *)p;
    }
    *out = g_scratch[dev];
    return 0;
}

}  // namespace

extern "C" int dengo_primordial_evolve_device_noalloc(
    primordial_data *data, const dengo_options *opt, double dtf, double dt0, double z,
    double *d_input, const double *d_rtol, double *d_atol, long long dims,
    const dengo_cell_outputs *outputs, double *ttot_min, double *dt_min, void *stream)
{
    if (!data || !d_input || !d_rtol || !d_atol) return set_error("null argument");
    if (dims < 0) return set_error("negative cell count");
    dengo_options o;
    if (check_options(opt, &o)) return -1;
    if (o.order)
        return set_error("the allocation-free entry point takes no cell order (its gather and "
                         "scatter buffers would need allocating)");
    dengo_cell_outputs out;
    memset(&out, 0, sizeof(out));
    if (outputs) out = *outputs;
    const bool per_cell_scratch = ttot_min && (!out.ttot || !out.dt);
    if (per_cell_scratch && dims > kMaxCells)
        return set_error("%lld cells: the allocation-free entry point holds per-cell elapsed "
                         "times for at most %lld cells (pass outputs->ttot and outputs->dt, or "
                         "no ttot_min, for more)", dims, kMaxCells);
    const NetHooks h = hooks_primordial();
    int full = 0;
    if (h.resident(data, &o, &full)) return -1;
    unsigned char *sp = nullptr;
    if (device_scratch(&sp)) return -1;
    double *ttot_cells = (double *)(sp + kHeader);
    double *dt_cells = ttot_cells + kMaxCells;
    double *red = (double *)(sp + 64);
    if (per_cell_scratch) {
        if (!out.ttot) out.ttot = ttot_cells;
        if (!out.dt) out.dt = dt_cells;
    }
    hipStream_t st = (hipStream_t)stream;
    LaunchArgs a;
    memset(&a, 0, sizeof(a));
    a.opt = o;
    if (a.opt.grid <= 0) a.opt.grid = full;
    a.dtf = dtf;
    a.dt0 = dt0;
    a.z = z;
    a.x = d_input;
    a.rtol = d_rtol;
    a.atol = d_atol;
    a.n = dims;
    a.out = out;
    a.stream = st;
    a.counter = (unsigned long long *)sp;   // reset by a kernel at the start of each launch
    a.ord = nullptr;
    data->current_z = z;
    if (h.launch(data, a)) return -1;
    if (ttot_min) {
        if (dims > 0) {
            hipLaunchKernelGGL(k_min_ttot<0>, dim3(1), dim3(1024), 0, st, out.ttot, out.dt, dims,
                               red);
            DENGO_HIP_OK(hipGetLastError());
            double r[2];
            DENGO_HIP_OK(hipMemcpyAsync(r, red, sizeof(r), hipMemcpyDeviceToHost, st));
            DENGO_HIP_OK(hipStreamSynchronize(st));
            *ttot_min = r[0];
            if (dt_min) *dt_min = r[1];
        } else {
            *ttot_min = 0.0;
            if (dt_min) *dt_min = dt0;
        }
    }
    return 0;
}
