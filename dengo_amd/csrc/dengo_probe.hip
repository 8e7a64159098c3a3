// Device FP64 vector rate: the compute ceiling of the lane-per-cell sweeps,
// measured on the device the caller has current, beside the published figure.
#include "dengo_solver.h"

namespace dengo {
namespace {
constexpr int kProbeChains = 16;   // independent FMA chains per lane
constexpr int kProbeThreads = 256;
constexpr int kProbeBlocksPerCU = 8;

// Each lane runs kProbeChains independent chains x <- x*a + b (v_fma_f64;
// a < 1 keeps x bounded and away from denormals) and writes their sum, so no
// chain is dead code.
__global__ void __launch_bounds__(kProbeThreads) k_fp64_fma_rate(double *out, int iters, double a,
                                                                 double b)
{
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double x[kProbeChains];
#pragma unroll
    for (int k = 0; k < kProbeChains; ++k) x[k] = 1.0 + 1e-3 * k + 1e-9 * (double)tid;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < kProbeChains; ++k) x[k] = __builtin_fma(x[k], a, b);
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kProbeChains; ++k) s += x[k];
    out[tid] = s;
}
}  // namespace
}  // namespace dengo

extern "C" int dengo_measure_fp64_fma_rate(int iters, double *tflops)
{
    using namespace dengo;
    if (!tflops) return set_error("tflops must not be NULL");
    if (iters <= 0 || iters > (1 << 24)) return set_error("iters must be in 1..2^24, got %d", iters);
    int dev = 0, cus = 0;
    DENGO_HIP_OK(hipGetDevice(&dev));
    DENGO_HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int blocks = cus * kProbeBlocksPerCU;
    const size_t threads = (size_t)blocks * kProbeThreads;
    double *out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    float ms = 0.0f;
    hipError_t e = hipMalloc(&out, threads * sizeof(double));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        // one untimed launch (code load, clocks up), then the timed one
        hipLaunchKernelGGL(k_fp64_fma_rate, dim3(blocks), dim3(kProbeThreads), 0, st, out,
                           iters < 64 ? iters : 64, 0.999999, 1e-6);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fp64_fma_rate, dim3(blocks), dim3(kProbeThreads), 0, st, out, iters,
                           0.999999, 1e-6);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) rc = set_error("fp64 rate probe: %s", hipGetErrorString(e));
    else if (ms <= 0.0f) rc = set_error("fp64 rate probe: no elapsed time");
    else *tflops = 2.0 * kProbeChains * (double)iters * (double)threads / (ms * 1e-3) / 1e12;
    if (e1) (void)hipEventDestroy(e1);
    if (e0) (void)hipEventDestroy(e0);
    if (st) (void)hipStreamDestroy(st);
    if (out) (void)hipFree(out);
    return rc;
}
