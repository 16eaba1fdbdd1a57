// peak_probe.hip -- peak-rate probes: the dense BF16 MFMA rate and the FP64
// FMA rate this GPU sustains, as reference points for the achieved rates the
// benchmark reports next to the published peaks.  Built into the diagnostics
// library (libbioshell_diag.so), not the product library.
#include <hip/hip_runtime.h>

#include <algorithm>

namespace bs {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Eight independent accumulators per wave: back-to-back v_mfma_f32_32x32x16_bf16.
__global__ __launch_bounds__(256) void probe_mfma_kernel(int iters, float* sink)
{
    const int lane = threadIdx.x & 63;
    bf16x8 a, b;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        a[k] = (__bf16)(1e-3f * (float)(lane + k));
        b[k] = (__bf16)(1e-3f * (float)(lane - k));
    }
    f32x16 acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[q][r] = 0.0f;
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[q], 0, 0, 0);
    float s = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; ++q)
#pragma unroll
        for (int r = 0; r < 16; ++r) s += acc[q][r];
    if (s == 1234.5f) sink[lane] = s;  // keeps the work; never true for these inputs
}

// Sixteen independent FP64 FMA chains per lane.
__global__ __launch_bounds__(256) void probe_fp64_kernel(int iters, float* sink)
{
    const int lane = threadIdx.x & 63;
    double x[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = 0.1 * (double)(lane + k);
    const double m = 0.999999, c = 1e-7;
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int k = 0; k < 16; ++k) x[k] = __builtin_fma(x[k], m, c);
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) s += x[k];
    if (s == 1234.5) sink[lane] = (float)s;
}

}  // namespace
}  // namespace bs

extern "C" {

// Sustained rate in TFLOP/s on the current device: which = 0 dense BF16 MFMA
// (32x32x16, FP32 accumulation), 1 FP64 vector FMA.  Best of three ~1 ms runs.
// Returns 0 on success, else a HIP error code.
int bsi_peak_probe(int which, double* tflops)
{
    using namespace bs;
    int dev = 0, n_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
    float* sink = nullptr;
    if (hipMalloc(&sink, 64 * sizeof(float)) != hipSuccess) return -1;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const int blocks = n_cu * (which == 0 ? 8 : 16);
    const int iters = which == 0 ? 512 : 2048;
    const double flops = which == 0 ? (double)blocks * 4 * iters * 8 * (32.0 * 32 * 16 * 2)
                                    : (double)blocks * 256 * iters * 16 * 2.0;
    float best = 1e30f;
    for (int rep = 0; rep < 4; ++rep) {
        (void)hipEventRecord(e0, nullptr);
        if (which == 0)
            probe_mfma_kernel<<<blocks, 256>>>(iters, sink);
        else
            probe_fp64_kernel<<<blocks, 256>>>(iters, sink);
        (void)hipEventRecord(e1, nullptr);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (rep > 0) best = std::min(best, ms);
    }
    const hipError_t err = hipGetLastError();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(sink);
    if (err != hipSuccess) return (int)err;
    *tflops = flops / (best * 1e-3) / 1e12;
    return 0;
}

}  // extern "C"
