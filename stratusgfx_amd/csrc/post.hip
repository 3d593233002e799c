// Consumers of the lit buffer.
//
// tonemapKernel — gammaTonemap.fs:125-142 (PerformGammaTonemapPostFx_,
// StratusRendererBackend.cpp:2356-2366): Narkowicz's ACES film fit
// saturate(c (a c + b) / (c (c' c + d) + e)) (:20-29), gamma correction
// pow(c, 1 / gamma) (:112-114), stored to the RGBA8 UNORM target (:674) as
// floor(c * 255 + 0.5), alpha 255.  One thread per pixel: 12 B read and 4 B
// written, HBM-bound.  Built with -ffp-contract=off, so the curve's products
// and sums round as the shader writes them; pow is the device library's.
#include "launch.h"

namespace {

__device__ __forceinline__ float sat(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

__device__ __forceinline__ uint32_t tonemapChannel(float x, float invGamma) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    const float film = sat((x * (a * x + b)) / (x * (c * x + d) + e));
    const float g = powf(film, invGamma);
    return (uint32_t)(sat(g) * 255.0f + 0.5f);
}

__global__ __launch_bounds__(256) void tonemapKernel(const float* __restrict__ lit, long n, float invGamma,
                                                      uint32_t* __restrict__ out) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float r = lit[3 * p], g = lit[3 * p + 1], b = lit[3 * p + 2];
    out[p] = tonemapChannel(r, invGamma) | (tonemapChannel(g, invGamma) << 8) | (tonemapChannel(b, invGamma) << 16) |
             0xFF000000u;
}

}  // namespace

namespace sgi_launch {

hipError_t tonemap(const float* lit, long n, float gamma, uint32_t* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const float invGamma = 1.0f / gamma;  // 1.0 / gamma (:113)
    hipLaunchKernelGGL(tonemapKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, lit, n, invGamma, out);
    return hipGetLastError();
}

}  // namespace sgi_launch
