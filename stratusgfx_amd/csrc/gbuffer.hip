// G-buffer unpacking: the reference's attachments (StratusRendererBackend.cpp:
// 384-424: RGB8 normals, RGBA8 albedo, RG8 base reflectivity, RGB8
// roughness / metallic / ambient, RG16F velocity) into the FP32 planes every
// pass reads.  A UNORM8 code k reads as k / 255 (correctly rounded, the GL
// conversion) and binary16 widens exactly.  One thread per pixel: 18 B read
// and 56 B written per pixel with velocity, an HBM-bound copy.
#include <hip/hip_fp16.h>

#include "launch.h"

namespace {

__device__ __forceinline__ float unorm8(uint32_t v, int byte) { return (float)((v >> (8 * byte)) & 0xFFu) / 255.0f; }

__global__ __launch_bounds__(256) void unpackKernel(sgi_gbuffer_packed g, long n, float* __restrict__ normal,
                                                     float* __restrict__ albedo, float* __restrict__ refl,
                                                     float* __restrict__ rma, float* __restrict__ vel) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t nv = g.normals[p], av = g.albedo[p], mv = g.roughnessMetallicAmbient[p];
    const uint32_t fv = g.baseReflectivity[p];
    for (int k = 0; k < 3; ++k) {
        normal[3 * p + k] = unorm8(nv, k);
        rma[3 * p + k] = unorm8(mv, k);
    }
    for (int k = 0; k < 4; ++k) albedo[4 * p + k] = unorm8(av, k);
    refl[2 * p] = unorm8(fv, 0);
    refl[2 * p + 1] = unorm8(fv, 1);
    if (vel) {
        const uint32_t vv = ((const uint32_t*)g.velocity)[p];
        vel[2 * p] = __half2float(__ushort_as_half((unsigned short)(vv & 0xFFFFu)));
        vel[2 * p + 1] = __half2float(__ushort_as_half((unsigned short)(vv >> 16)));
    }
}

}  // namespace

namespace sgi_launch {

hipError_t gbufferUnpack(const sgi_gbuffer_packed& g, float* normal, float* albedo, float* refl, float* rma,
                         float* velocity, hipStream_t s) {
    const long n = (long)g.width * g.height;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(unpackKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, n, normal, albedo, refl,
                       rma, velocity);
    return hipGetLastError();
}

}  // namespace sgi_launch
