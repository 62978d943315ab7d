// mti_probe.hip -- achievable-peak probes for the roofline figures: an HBM
// stream copy and an FP64 FMA loop, timed with HIP events.  A measurement
// helper loaded by the benchmark only; the imaging library does not use it.
#include <hip/hip_runtime.h>
#include <cstdint>

namespace {

// Grid-stride copy of n 16-byte elements.
__global__ __launch_bounds__(256) void k_copy16(const double2 *__restrict__ a, double2 *__restrict__ b, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

// Eight independent FP64 FMA chains per thread, iters steps each; one store
// per thread keeps the chains live.
__global__ __launch_bounds__(256) void k_fma64(double *out, int iters, double b, double c)
{
    double x[8];
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = 1e-3 * (threadIdx.x + k);
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int k = 0; k < 8; k++) x[k] = fma(x[k], b, c);
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 8; k++) s += x[k];
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

int fail(hipError_t e) { return e == hipSuccess ? 0 : (int)e; }

}  // namespace

// Device-to-device copy of `bytes` (rounded down to 16) repeated `reps`
// times; *gbps = (read + write bytes) / average time.  Returns a hipError_t.
extern "C" int mtp_copy_gbps(size_t bytes, int reps, double *gbps)
{
    const int64_t n = (int64_t)(bytes / 16);
    if (n <= 0 || reps <= 0 || !gbps) return (int)hipErrorInvalidValue;
    double2 *a = nullptr, *b = nullptr;
    hipEvent_t e0, e1;
    int r = fail(hipMalloc(&a, (size_t)n * 16));
    if (!r) r = fail(hipMalloc(&b, (size_t)n * 16));
    if (!r) r = fail(hipMemset(a, 0, (size_t)n * 16));
    if (!r) r = fail(hipEventCreate(&e0));
    if (!r) r = fail(hipEventCreate(&e1));
    if (!r) {
        const unsigned blocks = 8192;
        hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, 0, a, b, n);  // warm-up
        (void)hipEventRecord(e0, 0);
        for (int i = 0; i < reps; i++) hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, 0, a, b, n);
        (void)hipEventRecord(e1, 0);
        r = fail(hipEventSynchronize(e1));
        float ms = 0.f;
        if (!r) r = fail(hipEventElapsedTime(&ms, e0, e1));
        if (!r) *gbps = 2.0 * (double)n * 16.0 * reps / (ms * 1e-3) / 1e9;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return r;
}

// FP64 FMA rate: `blocks` x 256 threads x 8 chains x iters FMAs (2 flops
// each); *tflops from the average of `reps` launches.  Returns a hipError_t.
extern "C" int mtp_fp64_fma_tflops(int blocks, int iters, int reps, double *tflops)
{
    if (blocks <= 0 || iters <= 0 || reps <= 0 || !tflops) return (int)hipErrorInvalidValue;
    double *out = nullptr;
    hipEvent_t e0, e1;
    int r = fail(hipMalloc(&out, (size_t)blocks * 256 * sizeof(double)));
    if (!r) r = fail(hipEventCreate(&e0));
    if (!r) r = fail(hipEventCreate(&e1));
    if (!r) {
        hipLaunchKernelGGL(k_fma64, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999999, 1e-9);  // warm-up
        (void)hipEventRecord(e0, 0);
        for (int i = 0; i < reps; i++)
            hipLaunchKernelGGL(k_fma64, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999999, 1e-9);
        (void)hipEventRecord(e1, 0);
        r = fail(hipEventSynchronize(e1));
        float ms = 0.f;
        if (!r) r = fail(hipEventElapsedTime(&ms, e0, e1));
        if (!r) *tflops = 2.0 * 8.0 * (double)blocks * 256.0 * iters * reps / (ms * 1e-3) / 1e12;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    if (out) (void)hipFree(out);
    return r;
}
