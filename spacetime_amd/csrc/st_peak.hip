// Measured FP32 vector-FMA peak of the device: sixteen independent fmaf
// chains per lane over a grid of 2M lanes (the packed-FP32 forms are
// disabled in this build, so this is the non-packed issue rate every kernel
// here runs against).  bench.py reports it beside the datasheet peak.
#include <hip/hip_runtime.h>

#include "st_internal.h"

namespace st {
namespace {

constexpr int kChains = 16;
constexpr int kBlocks = 8192;
constexpr int kIters = 2048;

__global__ __launch_bounds__(256) void k_fma_peak(float* __restrict__ out, int iters, float b, float c) {
    float a[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) a[k] = (float)threadIdx.x * 1e-3f + (float)k;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) a[k] = fmaf(a[k], b, c);
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += a[k];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace
}  // namespace st

using st::fail;

extern "C" int st_measure_fp32_fma_peak(st_ctx* c, float* tflops) {
    if (!c || !tflops) return fail(ST_ERR_INVALID_ARG, "null argument");
    ST_HIP(hipSetDevice(c->device));
    float* buf = nullptr;
    if (hipMalloc(&buf, sizeof(float) * st::kBlocks * 256) != hipSuccess) return fail(ST_ERR_OOM, "hipMalloc failed");
    hipStream_t q = c->stream;
    hipLaunchKernelGGL(st::k_fma_peak, dim3(st::kBlocks), dim3(256), 0, q, buf, st::kIters, 0.999f, 1e-3f);
    const int reps = 3;
    hipError_t e = hipEventRecord(c->ev[0], q);
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        hipLaunchKernelGGL(st::k_fma_peak, dim3(st::kBlocks), dim3(256), 0, q, buf, st::kIters, 0.999f, 1e-3f);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(c->ev[1], q);
    if (e == hipSuccess) e = hipEventSynchronize(c->ev[1]);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(ST_ERR_DEVICE, hipGetErrorString(e));
    const double flops = 2.0 * reps * (double)st::kBlocks * 256.0 * st::kIters * st::kChains;
    *tflops = (float)(flops / (ms * 1e-3) / 1e12);
    c->have_last = false;
    return ST_OK;
}
