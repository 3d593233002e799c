// G-buffer unpacking: the reference's attachments (StratusRendererBackend.cpp:
// 384-424: RGB8 normals, RGBA8 albedo, RG8 base reflectivity, RGB8
// roughness / metallic / ambient, RG16F velocity) into the FP32 planes every
// pass reads.  A UNORM8 code k reads as k / 255 (correctly rounded, the GL
// conversion) and binary16 widens exactly.  One thread per pixel: 18 B read
// and 56 B written per pixel with velocity, an HBM-bound copy.
#include <hip/hip_fp16.h>

#include "launch.h"

namespace {

using sgi_dev::unorm8;  // the decode every pass's packed reads share

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

// The UNORM8 code GL stores for a fragment output v: clamp to [0, 1], then
// floor(v * 255 + 0.5) (NaN clamps to 0).
__device__ __forceinline__ uint32_t toUnorm8(float v) {
    const float c = v >= 0.0f ? (v <= 1.0f ? v : 1.0f) : 0.0f;
    return (uint32_t)floorf(__fadd_rn(__fmul_rn(c, 255.0f), 0.5f));
}

__global__ __launch_bounds__(256) void packKernel(const float* __restrict__ normal, const float* __restrict__ albedo,
                                                   const float* __restrict__ refl, const float* __restrict__ rma,
                                                   const float* __restrict__ vel, long n, uint32_t* __restrict__ n8,
                                                   uint32_t* __restrict__ a8, uint16_t* __restrict__ f8,
                                                   uint32_t* __restrict__ m8, uint32_t* __restrict__ v16) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    uint32_t nv = 0, mv = 0, av = 0;
    for (int k = 0; k < 3; ++k) {
        nv |= toUnorm8(normal[3 * p + k]) << (8 * k);
        mv |= toUnorm8(rma[3 * p + k]) << (8 * k);
    }
    for (int k = 0; k < 4; ++k) av |= toUnorm8(albedo[4 * p + k]) << (8 * k);
    n8[p] = nv;
    a8[p] = av;
    m8[p] = mv;
    f8[p] = (uint16_t)(toUnorm8(refl[2 * p]) | (toUnorm8(refl[2 * p + 1]) << 8));
    if (v16) {
        const uint32_t lo = __half_as_ushort(__float2half_rn(vel[2 * p]));
        const uint32_t hi = __half_as_ushort(__float2half_rn(vel[2 * p + 1]));
        v16[p] = lo | (hi << 16);
    }
}

}  // namespace

namespace sgi_launch {

hipError_t gbufferPack(const sgi_gbuffer& planes, const sgi_gbuffer_packed& out, hipStream_t s) {
    const long n = (long)planes.width * planes.height;
    if (n <= 0) return hipSuccess;
    uint32_t* v16 = (planes.velocity && out.velocity) ? (uint32_t*)const_cast<uint16_t*>(out.velocity) : nullptr;
    hipLaunchKernelGGL(packKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, planes.normal, planes.albedo,
                       planes.baseReflectivity, planes.roughnessMetallicAmbient, planes.velocity, n,
                       const_cast<uint32_t*>(out.normals), const_cast<uint32_t*>(out.albedo),
                       const_cast<uint16_t*>(out.baseReflectivity), const_cast<uint32_t*>(out.roughnessMetallicAmbient),
                       v16);
    return hipGetLastError();
}

hipError_t gbufferUnpack(const sgi_gbuffer_packed& g, float* normal, float* albedo, float* refl, float* rma,
                         float* velocity, hipStream_t s) {
    const long n = (long)g.width * g.height;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(unpackKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, n, normal, albedo, refl,
                       rma, velocity);
    return hipGetLastError();
}

}  // namespace sgi_launch
