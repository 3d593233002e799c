// Device helpers shared by the GI kernels: FP32 vector arithmetic in the
// reference shaders' evaluation order, GL texture addressing, the hash RNG and
// the cube / cascade lookups.
//
// The FP32 kernels are built with -ffp-contract=off and rely on HIP's default
// correctly rounded FP32 division and sqrt, so every decision the reference
// makes (side tests, radius tests, cube texel choice, sample counts, top-K
// order) is taken on the same bits as the FP32 restatement of the shaders.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sgi.h"

namespace sgi_dev {

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 splat(float s) { return f3{s, s, s}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 operator/(f3 a, f3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ f3 operator+(f3 a, float s) { return {a.x + s, a.y + s, a.z + s}; }
__device__ __forceinline__ f3 operator-(f3 a, float s) { return {a.x - s, a.y - s, a.z - s}; }
__device__ __forceinline__ f3 operator/(float s, f3 a) { return {s / a.x, s / a.y, s / a.z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float length(f3 a) { return __builtin_sqrtf(dot(a, a)); }
__device__ __forceinline__ f3 normalize(f3 a) { return a / length(a); }
// GLSL max / min as written in the restatement: a > b ? a : b.
__device__ __forceinline__ float gmax(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float gmin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float clampv(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
__device__ __forceinline__ float saturate(float x) { return clampv(x, 0.0f, 1.0f); }
__device__ __forceinline__ float mixv(float x, float y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ f3 mixv(f3 x, f3 y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ float pow5(float x) { const float x2 = x * x; return x2 * x2 * x; }
__device__ __forceinline__ float pow128(float x) {
#pragma unroll
    for (int i = 0; i < 7; ++i) x = x * x;
    return x;
}
__host__ __device__ __forceinline__ f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }

// Hardware reciprocal, reciprocal square root and exp2 (about 1 ulp).  Used
// only for continuous shading terms, never for a value a discrete decision of
// the reference depends on, so outputs stay within a few ulps of the exact
// restatement while every branch is taken on exactly rounded values.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fexp(float x) { return __builtin_amdgcn_exp2f(1.44269504088896f * x); }

constexpr float kPI = 3.14159265359f;     // common.glsl:3
constexpr float kEPS = 0.00001f;          // common.glsl:4 PREVENT_DIV_BY_ZERO
constexpr float kFLOAT_MAX = 3.402823466e+38f;

// ---- RNG (common.glsl:183-214) ----
__device__ __forceinline__ uint32_t hash1(uint32_t x) {
    x += (x << 10u);
    x ^= (x >> 6u);
    x += (x << 3u);
    x ^= (x >> 11u);
    x += (x << 15u);
    return x;
}
__device__ __forceinline__ float random3(float x, float y, float z) {
    uint32_t m = hash1(__float_as_uint(x) ^ hash1(__float_as_uint(y)) ^ hash1(__float_as_uint(z)));
    m &= 0x007FFFFFu;
    m |= 0x3F800000u;
    return __uint_as_float(m) - 1.0f;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
// floor(coordinate) -> texel index clamped to [0, n-1]; the float is clamped
// first so that NaN / huge coordinates convert without overflow (NaN -> 0).
__device__ __forceinline__ int texelIndex(float f, int n) {
    const float c = f > -1.0f ? (f < (float)n ? f : (float)n) : -1.0f;
    return clampi((int)c, 0, n - 1);
}

// Column-major mat4 * (x, y, z, w), summed column by column as GLSL does.
__device__ __forceinline__ void mulMat4(const float* m, float x, float y, float z, float w, float out[4]) {
#pragma unroll
    for (int r = 0; r < 4; ++r) out[r] = m[r] * x + m[4 + r] * y + m[8 + r] * z + m[12 + r] * w;
}

// common.glsl:143-155
__device__ __forceinline__ f3 worldPositionFromDepth(float u, float v, float depth, const float* invPV) {
    const float z = depth * 2.0f - 1.0f;
    float w[4];
    mulMat4(invPV, u * 2.0f - 1.0f, v * 2.0f - 1.0f, z, 1.0f, w);
    return f3{w[0] / w[3], w[1] / w[3], w[2] / w[3]};
}

__device__ __forceinline__ f3 decodeNormal(const float* p) {
    return normalize(f3{p[0], p[1], p[2]} * 2.0f - splat(1.0f));
}

// Texel-centre uv of pixel x in a viewport of size W.
__device__ __forceinline__ float texelCentre(int x, int W) { return ((float)x + 0.5f) / (float)W; }

// samplerCubeArray NEAREST (GL major-axis table; ties x, then y, then z).
// Returns the texel offset of direction d inside a layer's 6 faces.
__device__ __forceinline__ uint32_t cubeTexelOffset(f3 d, int res) {
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int face;
    float sc, tc, ma;
    if (ax >= ay && ax >= az) {
        ma = ax;
        if (d.x >= 0.0f) { face = 0; sc = -d.z; tc = -d.y; } else { face = 1; sc = d.z; tc = -d.y; }
    } else if (ay >= az) {
        ma = ay;
        if (d.y >= 0.0f) { face = 2; sc = d.x; tc = d.z; } else { face = 3; sc = d.x; tc = -d.z; }
    } else {
        ma = az;
        if (d.z >= 0.0f) { face = 4; sc = d.x; tc = -d.y; } else { face = 5; sc = -d.x; tc = -d.y; }
    }
    const float s = 0.5f * (sc / ma + 1.0f), t = 0.5f * (tc / ma + 1.0f);
    const int i = texelIndex(floorf(s * (float)res), res);
    const int j = texelIndex(floorf(t * (float)res), res);
    return ((uint32_t)face * res + j) * res + i;
}

// Kernel-argument view of the cube atlases.
struct CubeArgs {
    const void* atlas[SGI_MAX_TOTAL_SHADOW_ATLASES];
    int numAtlases;
    int res;
    int layers;
    uint32_t layerStride;  // 6 * res * res texels
};

// Base address of cube `layer` in atlas `a`; the unrolled select keeps the
// pointer table in kernel-argument registers instead of private memory.
__device__ __forceinline__ const void* cubeLayerBase(const CubeArgs& c, int a, int layer, int texelBytes) {
    const void* r = c.atlas[0];
    a = clampi(a, 0, c.numAtlases - 1);  // out-of-range entries read atlas 0
#pragma unroll
    for (int i = 1; i < SGI_MAX_TOTAL_SHADOW_ATLASES; ++i)
        if (a == i) r = c.atlas[i];
    const int l = clampi(layer, 0, c.layers - 1);
    return (const char*)r + (size_t)l * c.layerStride * texelBytes;
}

// floor(0.5 * (sc / ma + 1) * res) as the correctly rounded formula gives it,
// clamped to [0, res-1].  The estimate with a hardware reciprocal is within
// 5 * 2^-24 * res of the exact value, so when it lies more than res * 2^-19
// from an integer its floor is the exact floor; otherwise the exact formula
// runs (about 0.1 % of lookups at res 256).
// The rare exact paths are kept out of line so that the compiler does not
// if-convert them into every lane's common path.
__device__ __attribute__((noinline)) int cubeTexelCoordExact(float sc, float ma, int res) {
    return texelIndex(floorf(0.5f * (sc / ma + 1.0f) * (float)res), res);
}

__device__ __forceinline__ int cubeTexelCoord(float sc, float ma, int res) {
    const float fres = (float)res;
    const float x = 0.5f * (sc * frcp(ma) + 1.0f) * fres;
    const float fl = floorf(x);
    const float margin = fres * 0x1p-19f;
    if (__builtin_expect(x - fl > margin && (fl + 1.0f) - x > margin, 1)) return texelIndex(fl, res);
    return cubeTexelCoordExact(sc, ma, res);
}

// cubeTexelOffset with the exact floor decided from the estimate when safe;
// the face is chosen with selects (same table, same tie order).
__device__ __forceinline__ uint32_t cubeTexelOffsetFast(f3 d, int res) {
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    const bool xMajor = ax >= ay && ax >= az;
    const bool yMajor = !xMajor && ay >= az;
    const bool px = d.x >= 0.0f, py = d.y >= 0.0f, pz = d.z >= 0.0f;
    const float ma = xMajor ? ax : (yMajor ? ay : az);
    const int face = xMajor ? (px ? 0 : 1) : (yMajor ? (py ? 2 : 3) : (pz ? 4 : 5));
    const float sc = xMajor ? (px ? -d.z : d.z) : (yMajor ? d.x : (pz ? d.x : -d.x));
    const float tc = xMajor ? -d.y : (yMajor ? (py ? d.z : -d.z) : -d.y);
    const int i = cubeTexelCoord(sc, ma, res);
    const int j = cubeTexelCoord(tc, ma, res);
    return ((uint32_t)face * res + j) * res + i;
}

// True when dot(N, v) = a is far enough from zero, relative to the rounding
// error of both a and the reference's dot(N, v / len), for its sign to be the
// sign of the reference's value: a carries at most 3 * 2^-24 * sum|N_i v_i|,
// the reference at most 4 * 2^-24 * sum|N_i v_i| / len; the margin is 2^-20.
__device__ __forceinline__ bool signIsSafe(f3 N, f3 v, float a) {
    const float mag = fabsf(N.x * v.x) + fabsf(N.y * v.y) + fabsf(N.z * v.z);
    return mag > 1e-30f && fabsf(a) > mag * 0x1p-20f;
}

// Fused helpers for continuous terms only.
__device__ __forceinline__ float dotF(f3 a, f3 b) { return __builtin_fmaf(a.z, b.z, __builtin_fmaf(a.y, b.y, a.x * b.x)); }
__device__ __forceinline__ f3 fmaF(f3 a, float s, f3 c) {
    return f3{__builtin_fmaf(a.x, s, c.x), __builtin_fmaf(a.y, s, c.y), __builtin_fmaf(a.z, s, c.z)};
}

// Bytes per distance texel of a cube format.
__host__ __device__ constexpr int texelBytes(int fmt) { return fmt == SGI_TEXEL_F32 ? 4 : 2; }

// UNORM16 code k as k / 65535 correctly rounded: with y = k 2^-16 (exact),
// k / 65535 = y (1 + 2^-16 + 2^-32 + ...), and y + y (2^-16 + 2^-32) rounded
// once by the FMA equals the correctly rounded quotient for every k (checked
// exhaustively against the exact rationals by the tests).
__device__ __forceinline__ float unorm16ToFloat(uint32_t k) {
    const float y = (float)k * 0x1p-16f;
    return __builtin_fmaf(y, 0x1p-16f + 0x1p-32f, y);
}

// A 16-bit distance texel (F16 or UNORM16) as FP32.
template <int FMT>
__device__ __forceinline__ float decode16(unsigned short bits) {
    if constexpr (FMT == SGI_TEXEL_UNORM16) return unorm16ToFloat(bits);
    else return __half2float(__ushort_as_half(bits));
}

template <int FMT>
__device__ __forceinline__ float cubeFetch(const void* layerBase, int res, f3 d) {
    typedef const __attribute__((address_space(1))) unsigned short* GHalf;
    typedef const __attribute__((address_space(1))) float* GFloat;
    const uint32_t off = cubeTexelOffsetFast(d, res);
    if constexpr (FMT == SGI_TEXEL_F32) return ((GFloat)layerBase)[off];
    else return decode16<FMT>(((GHalf)layerBase)[off]);
}

// Kernel-argument view of the cascades and sun.
struct CsmArgs {
    const float* depth;
    int res;
    float projViews[SGI_NUM_CASCADES][16];
    float planes[SGI_NUM_CASCADES - 1][4];
    float bias;  // depthBias / (zfar - znear)
};

// sampler2DArrayShadow LINEAR + LEQUAL compare, CLAMP_TO_EDGE.
__device__ __forceinline__ float shadowCompare(const CsmArgs& c, float s, float t, int layer, float ref) {
    const float* base = c.depth + (size_t)layer * c.res * c.res;
    const float fu = s * (float)c.res - 0.5f, fv = t * (float)c.res - 0.5f;
    const float i0f = floorf(fu), j0f = floorf(fv);
    const float a = fu - i0f, b = fv - j0f;
    const int ia = texelIndex(i0f, c.res), ib = texelIndex(i0f + 1.0f, c.res);
    const int ja = texelIndex(j0f, c.res), jb = texelIndex(j0f + 1.0f, c.res);
    const float c00 = ref <= base[(size_t)ja * c.res + ia] ? 1.0f : 0.0f;
    const float c10 = ref <= base[(size_t)ja * c.res + ib] ? 1.0f : 0.0f;
    const float c01 = ref <= base[(size_t)jb * c.res + ia] ? 1.0f : 0.0f;
    const float c11 = ref <= base[(size_t)jb * c.res + ib] ? 1.0f : 0.0f;
    return (1.0f - a) * (1.0f - b) * c00 + a * (1.0f - b) * c10 + (1.0f - a) * b * c01 + a * b * c11;
}

__device__ __forceinline__ f3 cascadeBlends(const CsmArgs& c, f3 p) {
    float b[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) b[i] = c.planes[i][0] * p.x + c.planes[i][1] * p.y + c.planes[i][2] * p.z + c.planes[i][3] * 1.0f;
    return f3{b[0], b[1], b[2]};
}

// One cascade's 9 PCF taps (pbr.glsl:166-174), each the bilinear blend of 4
// LEQUAL compares, summed in the reference's (y, x) order.  The taps sit one
// texel apart, so together they cover a 4x4 texel block: the block is read and
// compared once, and each tap picks its 4 compares from it.  A tap whose
// rounded floor lands outside the block (its offset rounds across a texel
// edge) reads its own texels, so the result equals 9 shadowCompare calls.
__device__ __forceinline__ float pcf9Block(const CsmArgs& c, float s, float t, int layer, float ref) {
    const float* base = c.depth + (size_t)layer * c.res * c.res;
    const float fres = (float)c.res;
    const float wh = 1.0f / fres;
    const float I0 = floorf((s + -1.0f * wh) * fres - 0.5f);
    const float J0 = floorf((t + -1.0f * wh) * fres - 0.5f);
    uint32_t mask = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const size_t row = (size_t)texelIndex(J0 + (float)r, c.res) * c.res;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (ref <= base[row + texelIndex(I0 + (float)q, c.res)]) mask |= 1u << (r * 4 + q);
    }
    float light = 0.0f;
#pragma unroll
    for (int iy = -1; iy <= 1; ++iy) {
#pragma unroll
        for (int ix = -1; ix <= 1; ++ix) {
            const float x = (float)ix, y = (float)iy;
            const float fu = (s + x * wh) * fres - 0.5f, fv = (t + y * wh) * fres - 0.5f;
            const float i0f = floorf(fu), j0f = floorf(fv);
            const float a = fu - i0f, b = fv - j0f;
            const float dqf = i0f - I0, drf = j0f - J0;
            float c00, c10, c01, c11;
            if (dqf >= 0.0f && dqf <= 2.0f && drf >= 0.0f && drf <= 2.0f) {
                const int bit = (int)drf * 4 + (int)dqf;
                c00 = (mask >> bit) & 1u ? 1.0f : 0.0f;
                c10 = (mask >> (bit + 1)) & 1u ? 1.0f : 0.0f;
                c01 = (mask >> (bit + 4)) & 1u ? 1.0f : 0.0f;
                c11 = (mask >> (bit + 5)) & 1u ? 1.0f : 0.0f;
            } else {
                const int ia = texelIndex(i0f, c.res), ib = texelIndex(i0f + 1.0f, c.res);
                const int ja = texelIndex(j0f, c.res), jb = texelIndex(j0f + 1.0f, c.res);
                c00 = ref <= base[(size_t)ja * c.res + ia] ? 1.0f : 0.0f;
                c10 = ref <= base[(size_t)ja * c.res + ib] ? 1.0f : 0.0f;
                c01 = ref <= base[(size_t)jb * c.res + ia] ? 1.0f : 0.0f;
                c11 = ref <= base[(size_t)jb * c.res + ib] ? 1.0f : 0.0f;
            }
            light += (1.0f - a) * (1.0f - b) * c00 + a * (1.0f - b) * c10 + (1.0f - a) * b * c01 + a * b * c11;
        }
    }
    return light;
}

// pbr.glsl:158-227 calculateInfiniteShadowValue with the block-read PCF (the
// 9 taps of each of the two cascades from one 4x4 texel block: the same value
// as 36 single compares); `bias` already zero when unused.
__device__ __forceinline__ float infiniteShadowValueBlock(const CsmArgs& c, f3 sunDir, f3 fragPos, f3 blends, f3 normal,
                                                          float bias) {
    const f3 position = fragPos + normal * (1.0f - saturate(dot(normal, sunDir))) * 1.0f;
    const bool beyond2 = blends.y >= 0.0f, beyond3 = blends.z >= 0.0f;
    const int index1 = beyond2 ? 2 : 0, index2 = beyond3 ? 3 : 1;
    float c1[4], c2[4];
    mulMat4(c.projViews[index1], position.x, position.y, position.z, 1.0f, c1);
    mulMat4(c.projViews[index2], position.x, position.y, position.z, 1.0f, c2);
    const float s1 = (c1[0] / c1[3]) * 0.5f + 0.5f, t1 = (c1[1] / c1[3]) * 0.5f + 0.5f;
    const float depth1 = (c1[2] / c1[3]) * 0.5f + 0.5f;
    const float s2 = (c2[0] / c2[3]) * 0.5f + 0.5f, t2 = (c2[1] / c2[3]) * 0.5f + 0.5f;
    float depth2 = (c2[2] / c2[3]) * 0.5f + 0.5f;
    depth2 = beyond3 ? saturate(depth2) : depth2;
    const float weight = beyond2 ? saturate(blends.y) - saturate(blends.z) : 1.0f - saturate(blends.x);
    const float light1 = pcf9Block(c, s1, t1, index1, depth1 - bias);
    const float light2 = pcf9Block(c, s2, t2, index2, depth2 - bias);
    return mixv(light2, light1, weight) * (1.0f / 9.0f);
}

// Kernel-argument view of a G-buffer: the FP32 planes, or (n8 != nullptr)
// the reference's packed attachments for the normals, albedo, reflectivity,
// roughness / metallic / ambient and velocity (sgi_gbuffer.packed); depth, id
// and SSAO are FP32 either way.  The passes read it through the accessors
// below, which decode an attachment exactly as sgi_gbuffer_unpack does, so a
// pass gives the same bits on either form.
struct GbufArgs {
    int W, H;
    const float* depth;
    const float* normal;
    const float* albedo;
    const float* refl;
    const float* rma;
    const float* velocity;
    const float* id;
    const float* ssao;
    const uint32_t* n8;   // RGB8 normals (4-byte texels)
    const uint32_t* a8;   // RGBA8 albedo
    const uint16_t* f8;   // RG8 base reflectivity
    const uint32_t* m8;   // RGB8 roughness, metallic, ambient
    const uint32_t* v16;  // RG16F velocity (NULL: none)
};

__host__ inline GbufArgs gbufArgs(const sgi_gbuffer& g) {
    GbufArgs a{};
    a.W = g.width;
    a.H = g.height;
    a.depth = g.depth;
    a.id = g.id;
    a.ssao = g.ssao;
    if (g.packed) {
        a.n8 = g.packed->normals;
        a.a8 = g.packed->albedo;
        a.f8 = g.packed->baseReflectivity;
        a.m8 = g.packed->roughnessMetallicAmbient;
        a.v16 = reinterpret_cast<const uint32_t*>(g.packed->velocity);
    } else {
        a.normal = g.normal;
        a.albedo = g.albedo;
        a.refl = g.baseReflectivity;
        a.rma = g.roughnessMetallicAmbient;
        a.velocity = g.velocity;
    }
    return a;
}

// A UNORM8 channel (byte `byte` of a texel) as GL reads it: k / 255, correctly rounded.
__device__ __forceinline__ float unorm8(uint32_t v, int byte) { return (float)((v >> (8 * byte)) & 0xFFu) / 255.0f; }

// The stored normal of pixel p (n * 0.5 + 0.5 per channel, not yet decoded).
__device__ __forceinline__ f3 gNormalRaw(const GbufArgs& g, size_t p) {
    if (g.n8) {
        const uint32_t v = g.n8[p];
        return f3{unorm8(v, 0), unorm8(v, 1), unorm8(v, 2)};
    }
    return ld3(g.normal + 3 * p);
}
// The decoded normal (common.glsl's normal * 2 - 1, normalised) of pixel p.
__device__ __forceinline__ f3 gNormal(const GbufArgs& g, size_t p) {
    if (g.n8) {
        const uint32_t v = g.n8[p];
        return normalize(f3{unorm8(v, 0), unorm8(v, 1), unorm8(v, 2)} * 2.0f - splat(1.0f));
    }
    return decodeNormal(g.normal + 3 * p);
}
// Albedo rgb and emissive.r.
__device__ __forceinline__ float4 gAlbedo(const GbufArgs& g, size_t p) {
    if (g.a8) {
        const uint32_t v = g.a8[p];
        return make_float4(unorm8(v, 0), unorm8(v, 1), unorm8(v, 2), unorm8(v, 3));
    }
    const float* a = g.albedo + 4 * p;
    return make_float4(a[0], a[1], a[2], a[3]);
}
// Reflectance and emissive.g.
__device__ __forceinline__ float2 gRefl(const GbufArgs& g, size_t p) {
    if (g.f8) {
        const uint32_t v = g.f8[p];
        return make_float2(unorm8(v, 0), unorm8(v, 1));
    }
    return make_float2(g.refl[2 * p], g.refl[2 * p + 1]);
}
// Roughness, metallic and emissive.b.
__device__ __forceinline__ f3 gRma(const GbufArgs& g, size_t p) {
    if (g.m8) {
        const uint32_t v = g.m8[p];
        return f3{unorm8(v, 0), unorm8(v, 1), unorm8(v, 2)};
    }
    const float* m = g.rma + 3 * p;
    return f3{m[0], m[1], m[2]};
}
__device__ __forceinline__ bool gHasVelocity(const GbufArgs& g) { return g.n8 ? g.v16 != nullptr : g.velocity != nullptr; }
// The uv motion (cur - prev) of pixel p; 0 without a velocity plane.
__device__ __forceinline__ float2 gVelocity(const GbufArgs& g, size_t p) {
    if (g.n8) {
        if (!g.v16) return make_float2(0.0f, 0.0f);
        const uint32_t v = g.v16[p];
        return make_float2(__half2float(__ushort_as_half((unsigned short)(v & 0xFFFFu))),
                           __half2float(__ushort_as_half((unsigned short)(v >> 16))));
    }
    if (!g.velocity) return make_float2(0.0f, 0.0f);
    return make_float2(g.velocity[2 * p], g.velocity[2 * p + 1]);
}

// Per-frame uniforms the kernels read.
struct FrameArgs {
    float invPV[16];
    float viewPos[3];
    float sunDir[3];
    float sunColor[3];
    float time;
    float fps;
    float minRoughness;
    int perceptual;
    float minGiOcc;
    float fogDensity;
    float fogColor[3];
    float emissionStrength;
};

__host__ inline FrameArgs frameArgs(const sgi_frame& f) {
    FrameArgs a;
    for (int i = 0; i < 16; ++i) a.invPV[i] = f.invProjectionView[i];
    for (int k = 0; k < 3; ++k) {
        a.viewPos[k] = f.viewPosition[k];
        a.sunDir[k] = f.infiniteLightDirection[k];
        a.sunColor[k] = f.infiniteLightColor[k];
        a.fogColor[k] = f.fogColor[k];
    }
    a.time = f.time;
    a.fps = f.framesPerSecond;
    a.minRoughness = f.minRoughness;
    a.perceptual = f.usePerceptualRoughness;
    a.minGiOcc = f.minGiOcclusionFactor;
    a.fogDensity = f.fogDensity;
    a.emissionStrength = f.emissionStrength;
    return a;
}

// pbr2.glsl:139-144
__device__ __forceinline__ float remapRoughness(float r, const FrameArgs& f) {
    if (f.perceptual) r = r * r;
    return gmax(gmax(0.0f, f.minRoughness), r);
}
// pbr2.glsl:51-55
__device__ __forceinline__ float ndfGGX(float NdotH, float r) {
    const float r2 = r * r;
    const float f = (NdotH * r2 - NdotH) * NdotH + 1.0f;
    return r2 / gmax(kPI * f * f, kEPS);
}
// pbr2.glsl:59-64
__device__ __forceinline__ float visSmithGGX(float NdotV, float NdotL, float r) {
    const float a2 = r * r;
    const float ggxv = NdotL * __builtin_sqrtf((NdotV - a2 * NdotV) * NdotV + a2);
    const float ggxl = NdotV * __builtin_sqrtf((NdotL - a2 * NdotL) * NdotL + a2);
    return 0.5f / gmax(ggxv + ggxl, kEPS);
}
// pbr2.glsl:73-76
__device__ __forceinline__ f3 fresnelSchlick(float u, f3 f0) {
    const float f = pow5(1.0f - u);
    return splat(f) + f0 * (1.0f - f);
}
// pbr.glsl:260-267 with the length already known
__device__ __forceinline__ float vplAttenuation(float lightDist, float radius) {
    const float ratio = gmin(lightDist / radius, 1.0f);
    const float minDist = mixv(10.0f, 1.0f, expf(-3.0f * ratio)) * radius;
    return 1.0f / (minDist + 1.0f * lightDist * lightDist);
}

// XCD-aware block -> work mapping: hardware dispatches consecutive blocks round
// robin over the 8 XCDs, so hand each XCD a contiguous run of work items so
// that neighbouring tiles (which read the same VPL cube texels) share an L2.
__device__ __forceinline__ uint32_t xcdSwizzle(uint32_t b, uint32_t n) {
    constexpr uint32_t kXcd = 8;
    const uint32_t per = n / kXcd, rem = n % kXcd;
    const uint32_t x = b % kXcd, k = b / kXcd;
    // XCD x owns items [x*per + min(x, rem), ...) of length per + (x < rem)
    const uint32_t start = x * per + (x < rem ? x : rem);
    return start + k;
}

}  // namespace sgi_dev
