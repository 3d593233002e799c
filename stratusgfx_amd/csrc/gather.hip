// Per-pixel VPL gather — vpl_pbr_gi.fs:81-207.
//
// One 256-thread block shades one 16x16 screen tile (blocks are mapped to
// tiles XCD-contiguously so that neighbouring tiles, which fetch the same VPL
// cube texels, share an L2).  In tile-list mode the tile's list of up to K VPL
// records and atlas entries is staged once into LDS and every lane walks it
// with broadcast LDS reads; in the reference's random mode each pixel hashes
// its own VPL slots and reads the records through the cache hierarchy.
//
// Per-pixel quantities of the BRDF (fragment position, view vector, demodulated
// diffuse colour, f0, remapped roughness, N.V) are computed once per pixel.
// Everything a discrete decision of the reference depends on — the side and
// radius tests, the cube texel choice, the shadow compare, the sample count —
// is computed with correctly rounded division and sqrt in the restatement's
// operation order, so every branch matches the oracle bit for bit; the
// continuous BRDF terms use hardware rcp / rsq / exp2 (~1 ulp each).
#include "launch.h"
#include "views.h"

namespace {

using namespace sgi_dev;

struct Pixel {
    f3 P, V, N, f0, diffColor;
    float distToCamera, a, NdotV, ao, fog;
    float nvRaw;    // dot(N, V) before saturation
    float r2, r2m1; // remapped roughness^2 and roughness^2 - 1
    float gv;       // sqrt((NdotV - r2 NdotV) NdotV + r2), the view half of the Smith term
};

struct Vpl {  // the fields of sgi_vpl the gather reads, plus its cube's base address
    f3 pos, spec, color;
    float radius, invRadius, farPlane;
    const void* cube;
};

template <int FMT>
__device__ __forceinline__ Vpl loadVpl(const sgi_vpl* v, const sgi_atlas_entry* e, const CubeArgs& cubes) {
    Vpl o;
    o.pos = ld3(v->position);
    o.spec = ld3(v->specularPosition);
    o.color = ld3(v->color);
    o.radius = v->radius;
    o.invRadius = 1.0f / v->radius;
    o.farPlane = v->farPlane;
    const sgi_atlas_entry en = *e;
    o.cube = cubeLayerBase(cubes, en.index, en.layer, texelBytes(FMT));
    return o;
}

// A visible VPL's fields as the gather reads them, prepared once per frame for
// the sampled estimators (which fetch a random record per sample): 64 B, four
// 16-B loads, with the cube's base address already resolved.
struct PackedVpl {
    float4 posRadius;   // position, radius
    float4 specInvR;    // specular position, 1 / radius
    float4 colorFar;    // colour, farPlane
    const void* cube;
    uint64_t pad;
};

__device__ __forceinline__ PackedVpl packVpl(const Vpl& v) {
    PackedVpl o;
    o.posRadius = make_float4(v.pos.x, v.pos.y, v.pos.z, v.radius);
    o.specInvR = make_float4(v.spec.x, v.spec.y, v.spec.z, v.invRadius);
    o.colorFar = make_float4(v.color.x, v.color.y, v.color.z, v.farPlane);
    o.cube = v.cube;
    o.pad = 0;
    return o;
}

__device__ __forceinline__ Vpl unpackVpl(const PackedVpl& p) {
    Vpl v;
    v.pos = f3{p.posRadius.x, p.posRadius.y, p.posRadius.z};
    v.radius = p.posRadius.w;
    v.spec = f3{p.specInvR.x, p.specInvR.y, p.specInvR.z};
    v.invRadius = p.specInvR.w;
    v.color = f3{p.colorFar.x, p.colorFar.y, p.colorFar.z};
    v.farPlane = p.colorFar.w;
    v.cube = p.cube;
    return v;
}

// Beside the records the prep buffer holds each visible VPL's position and radius rounded to
// binary16 (8 B, SGI_MAX_TOTAL_VPLS_PER_FRAME of them after the records): the compacting gather's
// side / radius tests read this table, a quarter of the records' bytes, so that far more of its
// random reads hit the vector L1.
__host__ __device__ inline uint2* halfPositions(PackedVpl* prep) {
    return reinterpret_cast<uint2*>(prep + SGI_MAX_TOTAL_VPLS_PER_FRAME);
}

__device__ __forceinline__ uint32_t halfBits(float v) { return __half_as_ushort(__float2half(v)); }

__device__ __forceinline__ float halfValue(uint32_t bits) { return __half2float(__ushort_as_half((unsigned short)bits)); }

template <int FMT>
__global__ __launch_bounds__(256) void prepVplsKernel(const sgi_vpl* __restrict__ vpls,
                                                     const sgi_atlas_entry* __restrict__ entries,
                                                     const int32_t* __restrict__ visible, CubeArgs cubes,
                                                     PackedVpl* __restrict__ out) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int count = min(max(visible[0], 0), SGI_MAX_TOTAL_VPLS_PER_FRAME);
    if (s < count) {
        const PackedVpl r = packVpl(loadVpl<FMT>(vpls + s, entries + s, cubes));
        out[s] = r;
        const float4 q = r.posRadius;  // round to nearest even
        halfPositions(out)[s] = make_uint2(halfBits(q.x) | halfBits(q.y) << 16, halfBits(q.z) | halfBits(q.w) << 16);
    }
}

// vplDiffuseAttenuation (pbr.glsl:260-267) with the light distance known:
// 1 / (mix(10, 1, exp(-3 min(d / R, 1))) R + d^2).
__device__ __forceinline__ float vplAttenuationFast(float lightDist, float radius, float invRadius) {
    const float ratio = gmin(lightDist * invRadius, 1.0f);
    const float minDist = __builtin_fmaf(-9.0f, fexp(-3.0f * ratio), 10.0f) * radius;
    return frcp(__builtin_fmaf(lightDist, lightDist, minDist));
}

// The continuous part of one VPL's lighting, calculateVirtualPointLighting2 ->
// calculateLighting_Lambert_SeparateDiffuseSpecular (pbr2.glsl:541-574, 613-657),
// fused.  With unit V and Ls, H = (V + Ls) / |V + Ls| gives |V + Ls|^2 = 2 + 2 V.Ls,
// N.H = (N.V + N.Ls) / |V + Ls| and Ls.H = (1 + V.Ls) / |V + Ls|.
// nl = dot(N, L - P) (already computed for the side test).
__device__ __forceinline__ f3 vplLighting(const Pixel& px, const Vpl& L, float dist, float nl) {
    const f3 specDir = L.spec - px.P;
    const float s2 = dotF(specDir, specDir);
    const float sInv = frsq(s2);
    const float sLen = s2 * sInv;
    const f3 Ls = specDir * sInv;
    const float nlRaw = dotF(px.N, Ls);
    const float lv = dotF(Ls, px.V);
    const float hInv = frsq(__builtin_fmaf(2.0f, lv, 2.0f));
    const float NdotL = saturate(nlRaw);
    const float NdotH = saturate((px.nvRaw + nlRaw) * hInv);
    const float LdotH = saturate((1.0f + lv) * hInv);
    // NDF_GGX and Visibility_G_SmithGGX (pbr2.glsl:51-64)
    const float f = __builtin_fmaf(NdotH * px.r2m1, NdotH, 1.0f);
    const float ndf = px.r2 * frcp(gmax(kPI * f * f, kEPS));
    const float ggxl = px.NdotV * fsqrt(__builtin_fmaf(NdotL * -px.r2m1, NdotL, px.r2));
    const float vis = 0.5f * frcp(gmax(__builtin_fmaf(NdotL, px.gv, ggxl), kEPS));
    // Fr_Fresnel_Schlick (pbr2.glsl:73-76)
    const float fr5 = pow5(1.0f - LdotH);
    const f3 F = fmaF(px.f0, 1.0f - fr5, splat(fr5));
    const float spec = ndf * vis * NdotL * vplAttenuationFast(sLen, L.radius, L.invRadius);
    const float diff = saturate(nl * frcp(dist)) * (1.0f / kPI) * vplAttenuationFast(dist, L.radius, L.invRadius);
    const f3 brdf = fmaF(px.diffColor, diff, F * spec);
    // the ambient term is brdf * ao * colour * 0 and the fog intensity is 0, so
    // the lighting is (brdf * colour) * exp(-fogDensity * viewDist)
    return brdf * L.color;
}

// Exact side + radius test and exact shadow compare with the correctly rounded
// |L - P| (out of line: taken by a small fraction of lanes).
__device__ __attribute__((noinline)) bool rejectExact(f3 N, f3 lmf, float d2, float radius) {
    const float dist = __builtin_sqrtf(d2);
    return dot(N, lmf / dist) < 0.0f || dist > radius;
}

__device__ __attribute__((noinline)) bool shadowedExact(float d2, float m, float depth) {
    const float currentDepth = __builtin_sqrtf(d2);
    return (currentDepth - currentDepth * m) > depth;
}

// The visibility half of one VPL sample (vpl_pbr_gi.fs:149-199 up to the
// lighting call): the side and radius tests, the shadow-cube compare and the
// reweighting.  Returns false when the side / radius test rejects the VPL;
// otherwise adds the reweighting factor to validSamples and returns |L - P|,
// dot(N, L - P) and the lighting's scale (1 - shadow) * ssao * fog.
//
// |L - P| comes from the hardware sqrt (within a few ulps); each comparison
// that uses it is taken from that estimate only when it clears a margin of
// 2^-18 relative, far above the estimate's error, and otherwise from the
// correctly rounded length, so every branch matches the restatement.
struct Hit {
    float dist, nl, factor;
};

// The side and radius tests of vpl_pbr_gi.fs:163-170 (a VPL behind the surface
// or beyond its radius is rejected), on the decisions of the restatement.
// lmf = L - P, d2 = |lmf|^2, dist = the hardware sqrt of d2, nl = dot(N, lmf).
__device__ __forceinline__ bool passesSideRadius(const Pixel& px, f3 lmf, float d2, float dist, float nl,
                                                 float radius) {
    if (__builtin_expect(signIsSafe(px.N, lmf, nl) && fabsf(dist - radius) > radius * 0x1p-18f, 1))
        return !(nl < 0.0f || dist > radius);
    return !rejectExact(px.N, lmf, d2, radius);
}

// passesSideRadius's decision from the VPL's binary16 position and radius h (halfPositions):
// 1 accept, 0 reject, -1 when the rounding to binary16 could change it (the caller then runs
// passesSideRadius on the FP32 record).  A finite binary16 value v lies within 2^-11 |v| + 2^-25
// of the FP32 value it rounds, so over the three coordinates dot(N, L - P) and |L - P| move by at
// most E, the sum of those bounds (|N_i| <= 1), and the radius by its own bound.  A decision is
// taken only where it clears them by a further 2^-16 relative, far above the FP32 rounding of
// these sums and above the 2^-20 and 2^-18 margins within which passesSideRadius defers to the
// reference formula, so it is passesSideRadius's decision; an infinite or NaN value never clears.
__device__ __forceinline__ int sideRadiusHalf(const Pixel& px, uint2 h) {
    const float x = halfValue(h.x & 0xFFFFu), y = halfValue(h.x >> 16), z = halfValue(h.y & 0xFFFFu);
    const float r = halfValue(h.y >> 16);
    const f3 lmf = f3{x, y, z} - px.P;
    const float nl = dot(px.N, lmf);
    const float dist = fsqrt(dot(lmf, lmf));
    const float E = 0x1.01p-11f * (fabsf(x) + fabsf(y) + fabsf(z)) + 0x1p-23f;
    const float mag = fabsf(px.N.x * lmf.x) + fabsf(px.N.y * lmf.y) + fabsf(px.N.z * lmf.z) + E;
    const bool sideSure = fabsf(nl) > E + mag * 0x1p-16f;
    if (sideSure && nl < 0.0f) return 0;  // behind the surface, whatever the distance
    const float Er = E + 0x1.01p-11f * fabsf(r) + 0x1p-23f;
    const bool radiusSure = fabsf(dist - r) > Er + (dist + fabsf(r)) * 0x1p-16f;
    if (sideSure && radiusSure) return dist > r ? 0 : 1;
    return -1;
}

template <int FMT>
__device__ __forceinline__ bool visibility(const Pixel& px, const FrameArgs& fr, const CubeArgs& cubes, const Vpl& L,
                                           bool allowReject, float& validSamples, Hit& h) {
    const f3 lmf = L.pos - px.P;
    const float d2 = dot(lmf, lmf);
    const float dist = fsqrt(d2);
    const float nl = dot(px.N, lmf);  // exact order: side test and shadow bias
    if (allowReject && !passesSideRadius(px, lmf, d2, dist, nl, L.radius)) return false;
    const float distanceRatio = clampv((2.0f * dist) * L.invRadius, 0.0f, 1.0f);
    float shadowFactor = 0.0f;
    if (px.distToCamera < 500.0f) {  // calculateShadowValue1Sample, pbr.glsl:116-136
        const f3 fragDir = px.P - L.pos;
        // bias = |P - L| * max(minBias, saturate(dot(L - P, N)) * 0.03); dot(L - P, N)
        // has the same products in the same order as dot(N, L - P)
        const float m = gmax(0.1f, saturate(nl) * 0.03f);
        const float depth = cubeFetch<FMT>(L.cube, cubes.res, fragDir) * L.farPlane;
        const float lhs = dist - dist * m;
        bool shadowed;
        if (__builtin_expect(fabsf(lhs - depth) > dist * 0x1p-18f, 1))
            shadowed = lhs > depth;
        else
            shadowed = shadowedExact(d2, m, depth);
        shadowFactor = shadowed ? 1.0f : 0.0f;
    }
    // min(shadow, mix(minGiOcc, 1, ratio)) > 0 exactly when the shadow test fired
    shadowFactor = gmin(shadowFactor, mixv(fr.minGiOcc, 1.0f, distanceRatio));
    float reweightingFactor = 1.0f;
    if (shadowFactor > 0.0f) reweightingFactor = (1.0f - distanceRatio) * fr.minGiOcc + distanceRatio;
    validSamples += reweightingFactor;
    h.dist = dist;
    h.nl = nl;
    h.factor = (1.0f - shadowFactor) * px.ao * px.fog;
    return true;
}

// One VPL's whole contribution in FP32.  Returns false when rejected.
template <int FMT>
__device__ __forceinline__ bool evalVpl(const Pixel& px, const FrameArgs& fr, const CubeArgs& cubes, const Vpl& L,
                                        bool allowReject, f3& vplColor, float& validSamples) {
    Hit h;
    if (!visibility<FMT>(px, fr, cubes, L, allowReject, validSamples, h)) return false;
    vplColor = fmaF(vplLighting(px, L, h.dist, h.nl), h.factor, vplColor);
    return true;
}

// ---- packed-FP16 shading (SGI_PRECISION_FP16) ------------------------------
//
// The visibility half stays in FP32 with the exact decisions above, so the
// accepted samples, the shadow results and validSamples are those of the FP32
// path.  The lighting of two VPLs at a time runs in packed FP16 (one VPL per
// half of a register: v_pk_fma_f16 / v_pk_mul_f16 / v_pk_max_f16), and each
// product is accumulated into the FP32 colour with mixed-precision FMAs.
// Distances do not fit FP16 (|L - P|^2 reaches 2.5e5 and the attenuation falls
// to ~1e-6), so the attenuation is evaluated in a scaled form:
//   vplDiffuseAttenuation(d, R) = (64 / R^2) / (m(u) * 64 / R + u^2),  u = 8 d / R,
// with m = mix(10, 1, exp(-3 min(u / 8, 1))); the FP16 factor is at most R / 64
// and the 64 / R^2 goes into the FP32 colour scale.  The FP16 form therefore
// assumes VPL radii up to 16384 (the products stay below the FP16 maximum).
typedef _Float16 h1;
typedef _Float16 h2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ h2 hpair(float a, float b) { return h2{(h1)a, (h1)b}; }
__device__ __forceinline__ h2 hs(float a) { return h2{(h1)a, (h1)a}; }
__device__ __forceinline__ h2 hfma(h2 a, h2 b, h2 c) { return __builtin_elementwise_fma(a, b, c); }
__device__ __forceinline__ h2 hmax(h2 a, h2 b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ h2 hmin(h2 a, h2 b) { return __builtin_elementwise_min(a, b); }
__device__ __forceinline__ h2 hsat(h2 a) { return hmin(hmax(a, hs(0.0f)), hs(1.0f)); }
__device__ __forceinline__ h2 hrcp(h2 a) { return h2{(h1)__builtin_amdgcn_rcph(a.x), (h1)__builtin_amdgcn_rcph(a.y)}; }
__device__ __forceinline__ h2 hrsq(h2 a) { return h2{(h1)__builtin_amdgcn_rsqh(a.x), (h1)__builtin_amdgcn_rsqh(a.y)}; }
__device__ __forceinline__ h2 hsqrt(h2 a) { return __builtin_elementwise_sqrt(a); }
__device__ __forceinline__ h2 hexp2(h2 a) { return __builtin_elementwise_exp2(a); }

struct PixelH {  // the pixel's shading inputs, both halves equal
    h2 ndotv, r2, r2m1, gv, halfInvGv;
    h2 f0r, f0g, f0b, dr, dg, db;
    bool nvZero;  // N.V <= 0 (saturated to 0)
};

// The three dot products of the unit specular direction Ls are formed in FP32:
// rounded to FP16 first, 1 + V.Ls cancels when the light lies opposite the
// viewer (V.Ls near -1, a grazing view) and the half vector's length and
// direction lose every significant bit.
struct LightH {  // one VPL's FP32 inputs to the packed stage
    float nl;          // N.Ls
    float opl;         // 1 + V.Ls = |V + Ls|^2 / 2
    float nvl;         // N.V + N.Ls = N.(V + Ls)
    float uD, uS;      // 8 |L - P| / R and 8 |S - P| / R
    float cosD;        // dot(N, L - P) / |L - P|
    float invR64;      // 64 / R
    f3 col;            // colour * (1 - shadow) * ssao * fog * 64 / R^2 (0 for no light)
};

__device__ __forceinline__ LightH noLightH() {
    LightH o;
    o.nl = 0.0f; o.opl = 1.0f; o.nvl = 0.0f;
    o.uD = 0.0f; o.uS = 0.0f; o.cosD = 0.0f; o.invR64 = 1.0f;
    o.col = splat(0.0f);
    return o;
}

__device__ __forceinline__ PixelH toHalf(const Pixel& px) {
    PixelH q;
    q.ndotv = hs(px.NdotV); q.r2 = hs(px.r2); q.r2m1 = hs(px.r2m1); q.gv = hs(px.gv);
    q.halfInvGv = hs(0.5f / px.gv);
    q.nvZero = px.NdotV == 0.0f;
    q.f0r = hs(px.f0.x); q.f0g = hs(px.f0.y); q.f0b = hs(px.f0.z);
    q.dr = hs(px.diffColor.x); q.dg = hs(px.diffColor.y); q.db = hs(px.diffColor.z);
    return q;
}

__device__ __forceinline__ LightH prepHalf(const Pixel& px, const Vpl& L, const Hit& h) {
    const f3 specDir = L.spec - px.P;
    const float s2 = dotF(specDir, specDir);
    const float sInv = frsq(s2);
    LightH o;
    o.nl = dotF(px.N, specDir) * sInv;
    o.opl = 1.0f + dotF(px.V, specDir) * sInv;
    o.nvl = px.nvRaw + o.nl;
    const float invR8 = 8.0f * L.invRadius;
    o.uD = h.dist * invR8;
    o.uS = gmin((s2 * sInv) * invR8, 240.0f);
    o.cosD = h.nl * frcp(h.dist);
    o.invR64 = 64.0f * L.invRadius;
    o.col = L.color * (h.factor * L.invRadius * o.invR64);
    return o;
}

// The lighting of two VPLs (pbr2.glsl:51-76, 541-574, 613-657 and pbr.glsl:260-267,
// rearranged as in vplLighting), accumulated into acc.
__device__ __forceinline__ void shadePairH(const PixelH& q, const LightH& a, const LightH& b, f3& acc) {
    const h2 nlRaw = hpair(a.nl, b.nl), opl = hpair(a.opl, b.opl), nvl = hpair(a.nvl, b.nvl);
    // |V + Ls|^2 = 2 (1 + V.Ls), kept >= 2^-14 so that H stays finite
    const h2 hInv = hrsq(hmax(hs(2.0f) * opl, hs(0x1p-14f)));
    const h2 NdotL = hsat(nlRaw);
    const h2 NdotH = hsat(nvl * hInv);
    const h2 LdotH = hsat(opl * hInv);
    const h2 f = hfma(NdotH * q.r2m1, NdotH, hs(1.0f));
    // the FP32 path's 1e-5 floors are below FP16's normal range (subnormals flush), so
    // the floors here are 2^-14, the smallest normal FP16 value
    const h2 ndf = q.r2 * hrcp(hmax(hs(kPI) * f * f, hs(0x1p-14f)));
    const h2 ggxl = q.ndotv * hsqrt(hfma(NdotL * -q.r2m1, NdotL, q.r2));
    // Vis * N.L <= 2: the product is formed before it meets the NDF.  At N.V = 0
    // it is 0.5 / gv for any N.L > 0 (and 0 at N.L = 0), which FP16 cannot reach
    // through the quotient once N.L * gv falls below the normal range.
    const h2 visNl = q.nvZero ? hmin(NdotL * hs(16384.0f), hs(1.0f)) * q.halfInvGv
                              : (hs(0.5f) * NdotL) * hrcp(hmax(hfma(NdotL, q.gv, ggxl), hs(0x1p-14f)));
    const h2 specT = ndf * visNl;
    const h2 t = hs(1.0f) - LdotH;
    const h2 t2 = t * t;
    const h2 fr5 = t2 * t2 * t;
    const h2 omf = hs(1.0f) - fr5;
    const h2 uD = hpair(a.uD, b.uD), uS = hpair(a.uS, b.uS), iR = hpair(a.invR64, b.invR64);
    const h2 kE = hs(-3.0f * 1.44269504088896f / 8.0f);
    const h2 mD = hfma(hs(-9.0f), hexp2(hmin(uD, hs(8.0f)) * kE), hs(10.0f));
    const h2 mS = hfma(hs(-9.0f), hexp2(hmin(uS, hs(8.0f)) * kE), hs(10.0f));
    const h2 AD = hrcp(hfma(mD, iR, uD * uD));
    const h2 AS = hrcp(hfma(mS, iR, uS * uS));
    const h2 diff = hsat(hpair(a.cosD, b.cosD)) * hs(1.0f / kPI) * AD;
    const h2 spec = specT * AS;
    const h2 wr = hfma(q.dr, diff, hfma(q.f0r, omf, fr5) * spec);
    const h2 wg = hfma(q.dg, diff, hfma(q.f0g, omf, fr5) * spec);
    const h2 wb = hfma(q.db, diff, hfma(q.f0b, omf, fr5) * spec);
    acc.x = __builtin_fmaf((float)wr.x, a.col.x, acc.x);
    acc.y = __builtin_fmaf((float)wg.x, a.col.y, acc.y);
    acc.z = __builtin_fmaf((float)wb.x, a.col.z, acc.z);
    acc.x = __builtin_fmaf((float)wr.y, b.col.x, acc.x);
    acc.y = __builtin_fmaf((float)wg.y, b.col.y, acc.y);
    acc.z = __builtin_fmaf((float)wb.y, b.col.z, acc.z);
}

__device__ __forceinline__ Pixel setupPixel(const GbufArgs& g, const FrameArgs& fr, int x, int y, size_t p) {
    Pixel px;
    px.P = worldPositionFromDepth(texelCentre(x, g.W), texelCentre(y, g.H), g.depth[p], fr.invPV);
    const f3 vp = ld3(fr.viewPos) - px.P;
    const float vp2 = dot(vp, vp);
    px.distToCamera = __builtin_sqrtf(vp2);  // exact: shadow on / off and the sample count
    px.V = vp * frsq(vp2);
    const float4 al = gAlbedo(g, p);
    const f3 baseColor = f3{gmax(al.x, 0.01f), gmax(al.y, 0.01f), gmax(al.z, 0.01f)};
    px.N = gNormal(g, p);  // exact: side test
    const f3 rma = gRma(g, p);
    const float roughness = gmax(0.5f, rma.x);
    const float metallic = rma.y;
    px.ao = clampv(g.ssao ? g.ssao[p] : 1.0f, 0.35f, 1.0f);
    const f3 baseRefl = splat(gRefl(g, p).x);
    // BRDF_Base with diffuseDivisor = baseColor, specularMultiplier = 1 / (baseColor + eps)
    px.a = remapRoughness(roughness, fr);
    const f3 specMul = f3{frcp(baseColor.x + kEPS), frcp(baseColor.y + kEPS), frcp(baseColor.z + kEPS)};
    px.diffColor = ((1.0f - metallic) * baseColor) * specMul;
    px.f0 = (0.16f * baseRefl * baseRefl * (1.0f - metallic) + baseColor * metallic) * specMul;
    px.nvRaw = dot(px.N, px.V);
    px.NdotV = saturate(px.nvRaw);
    px.r2 = px.a * px.a;
    px.r2m1 = px.r2 - 1.0f;
    px.gv = fsqrt((px.NdotV - px.r2 * px.NdotV) * px.NdotV + px.r2);
    px.fog = fexp(-fr.fogDensity * px.distToCamera);
    return px;
}

__device__ __forceinline__ void writeOut(const GbufArgs& g, size_t p, f3 vplColor, float validSamples, float* color,
                                         float* reservoir) {
    const float4 al = gAlbedo(g, p);
    color[3 * p + 0] = gmax(al.x, 0.01f) + kEPS;
    color[3 * p + 1] = gmax(al.y, 0.01f) + kEPS;
    color[3 * p + 2] = gmax(al.z, 0.01f) + kEPS;
    float4 r;
    r.x = gmin(vplColor.x, 65504.0f);
    r.y = gmin(vplColor.y, 65504.0f);
    r.z = gmin(vplColor.z, 65504.0f);
    r.w = gmax(validSamples, 1.0f);
    reinterpret_cast<float4*>(reservoir)[p] = r;
}

constexpr int kMaxK = SGI_MAX_VPLS_PER_TILE_LIMIT;

template <int FMT, int TS, bool HALF>
__device__ __forceinline__ void gatherTileBody(const GbufArgs& g, const FrameArgs& fr, const CubeArgs& cubes, int t,
                                               int tilesX, int K, const sgi_vpl* __restrict__ vpls,
                                               const sgi_atlas_entry* __restrict__ entries,
                                               const int32_t* __restrict__ tiles, float* __restrict__ color,
                                               float* __restrict__ reservoir, Vpl* lv, int& ln) {
    const int tx = t % tilesX, ty = t / tilesX;
    const int32_t* rec = tiles + (size_t)t * (1 + K);
    if (threadIdx.x == 0) ln = min(max(rec[0], 0), K);
    if ((int)threadIdx.x < K) {
        const int n = rec[0];
        if ((int)threadIdx.x < n) {
            const int s = rec[1 + threadIdx.x];
            lv[threadIdx.x] = loadVpl<FMT>(vpls + s, entries + s, cubes);
        }
    }
    __syncthreads();
    const int x = tx * TS + (int)threadIdx.x % TS, y = ty * TS + (int)threadIdx.x / TS;
    if (x >= g.W || y >= g.H) return;
    const size_t p = (size_t)y * g.W + x;
    const Pixel px = setupPixel(g, fr, x, y, p);
    f3 vplColor = splat(0.0f);
    float validSamples = 0.0f;
    const int n = ln;
    if constexpr (HALF) {
        // list positions k and k + 1 share one packed evaluation
        const PixelH q = toHalf(px);
        for (int k = 0; k < n; k += 2) {
            LightH a = noLightH(), b = noLightH();
            Hit h;
            if (visibility<FMT>(px, fr, cubes, lv[k], true, validSamples, h)) a = prepHalf(px, lv[k], h);
            if (k + 1 < n && visibility<FMT>(px, fr, cubes, lv[k + 1], true, validSamples, h))
                b = prepHalf(px, lv[k + 1], h);
            shadePairH(q, a, b, vplColor);
        }
    } else {
        for (int k = 0; k < n; ++k) evalVpl<FMT>(px, fr, cubes, lv[k], true, vplColor, validSamples);
    }
    writeOut(g, p, vplColor, validSamples, color, reservoir);
}

template <int FMT, int TS, bool HALF>
__global__ __launch_bounds__(TS * TS) void gatherTilesKernel(GbufArgs g, FrameArgs fr, CubeArgs cubes, int tilesX,
                                                            int nTiles, int K, const sgi_vpl* __restrict__ vpls,
                                                            const sgi_atlas_entry* __restrict__ entries,
                                                            const int32_t* __restrict__ tiles,
                                                            float* __restrict__ color, float* __restrict__ reservoir) {
    __shared__ Vpl lv[kMaxK];
    __shared__ int ln;
    gatherTileBody<FMT, TS, HALF>(g, fr, cubes, (int)xcdSwizzle(blockIdx.x, (uint32_t)nTiles), tilesX, K, vpls,
                                  entries, tiles, color, reservoir, lv, ln);
}

// Batched: the tiles of every view, numbered view after view (FP32 shading).
template <int FMT, int TS>
__global__ __launch_bounds__(TS * TS) void gatherTilesViewsKernel(const sgi_batch::ViewDesc* __restrict__ views,
                                                                 int numViews, int totalTiles, int K, CubeArgs cubes) {
    __shared__ Vpl lv[kMaxK];
    __shared__ int ln;
    const int gt = (int)xcdSwizzle(blockIdx.x, (uint32_t)totalTiles);
    const sgi_batch::ViewDesc& d = views[sgi_batch::findView<sgi_batch::kTiles>(views, numViews, gt)];
    gatherTileBody<FMT, TS, false>(d.g, d.fr, cubes, gt - d.tileStart, d.tilesX, K, d.vplUpdated, d.entries, d.tiles,
                                   d.color, d.reservoir, lv, ln);
}

// vpl_pbr_gi.fs:158-161: slot int(maxRandomIndex * random(fragCoord, seedZ)),
// then the seed moves on by 10000.
__device__ __forceinline__ int drawSlot(float seedX, float seedY, float& seedZ, int maxRandomIndex) {
    const float rnd = random3(seedX, seedY, seedZ);
    seedZ += 10000.0f;
    return (int)((float)maxRandomIndex * rnd);
}

// The same draw over every visible slot (the last one included).
__device__ __forceinline__ int drawSlotAll(float seedX, float seedY, float& seedZ, int count) {
    const float rnd = random3(seedX, seedY, seedZ);
    seedZ += 10000.0f;
    return min((int)((float)count * rnd), count - 1);
}

// The sampled estimators draw one hashed slot per loop iteration whether the
// sample is accepted or rejected, so the slot of iteration k + 1 is known before
// iteration k is shaded: its record is fetched one iteration ahead.
template <int FMT, bool HALF>
__global__ __launch_bounds__(256) void gatherRandomKernel(GbufArgs g, FrameArgs fr, CubeArgs cubes, int tilesX,
                                                         int nTiles, const PackedVpl* __restrict__ prep,
                                                         const int32_t* __restrict__ visible,
                                                         float* __restrict__ color, float* __restrict__ reservoir) {
    const uint32_t t = xcdSwizzle(blockIdx.x, (uint32_t)nTiles);
    const int tx = (int)t % tilesX, ty = (int)t / tilesX;
    const int x = tx * 16 + (int)threadIdx.x % 16, y = ty * 16 + (int)threadIdx.x / 16;
    if (x >= g.W || y >= g.H) return;
    const size_t p = (size_t)y * g.W + x;
    const Pixel px = setupPixel(g, fr, x, y, p);
    f3 vplColor = splat(0.0f);
    float validSamples = 0.0f;
    // prep holds min(visible[0], 4480) records (prepVplsKernel): a larger
    // caller-supplied count must not index past them
    const int numVisible = min(max(visible[0], 0), SGI_MAX_TOTAL_VPLS_PER_FRAME);
    if (numVisible > 0) {
        const float roughness = gmax(0.5f, gRma(g, p).x);
        const float roughnessWeight = 1.0f - roughness;
        const float distRatioToCamera = gmin(1.0f - px.distToCamera / 1000.0f, 1.0f);
        const int maxSamplesPerPixel = (int)mixv(2.0f, 10.0f, roughnessWeight);
        int samplesMax = (int)((float)maxSamplesPerPixel * distRatioToCamera);
        samplesMax = samplesMax > 1 ? samplesMax : 1;
        const int maxRandomIndex = numVisible - 1;
        const float seedX = (float)x + 0.5f, seedY = (float)y + 0.5f;
        float seedZ = fr.time;
        if constexpr (HALF) {
            // accepted samples are paired in the order drawn
            const PixelH q = toHalf(px);
            LightH pending = noLightH();
            bool hasPending = false;
            PackedVpl next = prep[drawSlot(seedX, seedY, seedZ, maxRandomIndex)];
            for (int i = 0, resamples = 0; i < samplesMax; i += 1) {
                const Vpl L = unpackVpl(next);
                next = prep[drawSlot(seedX, seedY, seedZ, maxRandomIndex)];
                Hit h;
                if (!visibility<FMT>(px, fr, cubes, L, resamples < 5, validSamples, h)) {
                    ++resamples;
                    --i;
                    continue;
                }
                const LightH cur = prepHalf(px, L, h);
                if (hasPending) shadePairH(q, pending, cur, vplColor);
                else pending = cur;
                hasPending = !hasPending;
            }
            if (hasPending) shadePairH(q, pending, noLightH(), vplColor);
        } else {
            PackedVpl next = prep[drawSlot(seedX, seedY, seedZ, maxRandomIndex)];
            for (int i = 0, resamples = 0; i < samplesMax; i += 1) {
                const Vpl L = unpackVpl(next);
                next = prep[drawSlot(seedX, seedY, seedZ, maxRandomIndex)];
                if (!evalVpl<FMT>(px, fr, cubes, L, resamples < 5, vplColor, validSamples)) {
                    ++resamples;
                    --i;
                    continue;
                }
            }
        }
    }
    writeOut(g, p, vplColor, validSamples, color, reservoir);
}

// Every visible VPL per pixel (the converged estimate), optionally summed in
// FP64.  The VPLs are staged through LDS 256 at a time and every lane of the
// 16x16 tile walks the chunk.
template <int FMT, bool ACC64>
__global__ __launch_bounds__(256) void gatherAllKernel(GbufArgs g, FrameArgs fr, CubeArgs cubes, int tilesX,
                                                      int nTiles, const sgi_vpl* __restrict__ vpls,
                                                      const int32_t* __restrict__ visible,
                                                      const sgi_atlas_entry* __restrict__ entries,
                                                      float* __restrict__ color, float* __restrict__ reservoir) {
    __shared__ Vpl lv[256];
    const uint32_t t = xcdSwizzle(blockIdx.x, (uint32_t)nTiles);
    const int tx = (int)t % tilesX, ty = (int)t / tilesX;
    const int x = tx * 16 + (int)threadIdx.x % 16, y = ty * 16 + (int)threadIdx.x / 16;
    const bool inside = x < g.W && y < g.H;
    const size_t p = inside ? (size_t)y * g.W + x : 0;
    Pixel px;
    if (inside) px = setupPixel(g, fr, x, y, p);
    const int count = min(max(visible[0], 0), SGI_MAX_TOTAL_VPLS_PER_FRAME);
    double c64[4] = {0.0, 0.0, 0.0, 0.0};
    f3 vplColor = splat(0.0f);
    float validSamples = 0.0f;
    for (int base = 0; base < count; base += 256) {
        const int n = min(256, count - base);
        __syncthreads();
        if ((int)threadIdx.x < n) lv[threadIdx.x] = loadVpl<FMT>(vpls + base + threadIdx.x, entries + base + threadIdx.x, cubes);
        __syncthreads();
        if (!inside) continue;
        for (int k = 0; k < n; ++k) {
            if (ACC64) {
                f3 c = splat(0.0f);
                float w = 0.0f;
                if (evalVpl<FMT>(px, fr, cubes, lv[k], true, c, w)) {
                    c64[0] += (double)c.x;
                    c64[1] += (double)c.y;
                    c64[2] += (double)c.z;
                    c64[3] += (double)w;
                }
            } else {
                evalVpl<FMT>(px, fr, cubes, lv[k], true, vplColor, validSamples);
            }
        }
    }
    if (!inside) return;
    if (ACC64) {
        vplColor = f3{(float)c64[0], (float)c64[1], (float)c64[2]};
        validSamples = (float)c64[3];
    }
    writeOut(g, p, vplColor, validSamples, color, reservoir);
}

// `samples` hashed-random slots per pixel over all visible VPLs, with the
// reference's hash and seed stepping; the rejection rule as sgi_gather_opts.
// HALF (fixed draw count only, rejectionLimit < 0): the draws are shaded two
// at a time in packed FP16 (visibility and acceptance stay FP32), in the same
// draw order.
template <int FMT, bool HALF>
__global__ __launch_bounds__(256) void gatherSamplesKernel(GbufArgs g, FrameArgs fr, CubeArgs cubes, int tilesX,
                                                          int nTiles, int samples, int rejectionLimit,
                                                          const PackedVpl* __restrict__ prep,
                                                          const int32_t* __restrict__ visible,
                                                          float* __restrict__ color, float* __restrict__ reservoir) {
    const uint32_t t = xcdSwizzle(blockIdx.x, (uint32_t)nTiles);
    const int tx = (int)t % tilesX, ty = (int)t / tilesX;
    const int x = tx * 16 + (int)threadIdx.x % 16, y = ty * 16 + (int)threadIdx.x / 16;
    if (x >= g.W || y >= g.H) return;
    const size_t p = (size_t)y * g.W + x;
    const Pixel px = setupPixel(g, fr, x, y, p);
    f3 vplColor = splat(0.0f);
    float validSamples = 0.0f;
    const int count = min(max(visible[0], 0), SGI_MAX_TOTAL_VPLS_PER_FRAME);
    if (count > 0) {
        const float seedX = (float)x + 0.5f, seedY = (float)y + 0.5f;
        float seedZ = fr.time;
        PackedVpl next = prep[drawSlotAll(seedX, seedY, seedZ, count)];
        if constexpr (HALF) {
            const PixelH q = toHalf(px);
            PackedVpl nextB = prep[drawSlotAll(seedX, seedY, seedZ, count)];
            int i = 0;
            for (; i + 1 < samples; i += 2) {
                const Vpl La = unpackVpl(next), Lb = unpackVpl(nextB);
                next = prep[drawSlotAll(seedX, seedY, seedZ, count)];
                nextB = prep[drawSlotAll(seedX, seedY, seedZ, count)];
                LightH a = noLightH(), b = noLightH();
                Hit h;
                if (visibility<FMT>(px, fr, cubes, La, true, validSamples, h)) a = prepHalf(px, La, h);
                if (visibility<FMT>(px, fr, cubes, Lb, true, validSamples, h)) b = prepHalf(px, Lb, h);
                shadePairH(q, a, b, vplColor);
            }
            if (i < samples) {  // an odd count: the last draw alone
                const Vpl La = unpackVpl(next);
                LightH a = noLightH();
                Hit h;
                if (visibility<FMT>(px, fr, cubes, La, true, validSamples, h)) a = prepHalf(px, La, h);
                shadePairH(q, a, noLightH(), vplColor);
            }
            writeOut(g, p, vplColor, validSamples, color, reservoir);
            return;
        }
        for (int i = 0, resamples = 0; i < samples; i += 1) {
            const Vpl L = unpackVpl(next);
            next = prep[drawSlotAll(seedX, seedY, seedZ, count)];
            if (rejectionLimit < 0) {  // fixed draw count: a failing draw is skipped
                evalVpl<FMT>(px, fr, cubes, L, true, vplColor, validSamples);
                continue;
            }
            if (!evalVpl<FMT>(px, fr, cubes, L, resamples < rejectionLimit, vplColor, validSamples)) {
                ++resamples;
                --i;
                continue;
            }
        }
    }
    writeOut(g, p, vplColor, validSamples, color, reservoir);
}

// The fixed-count sampler (rejectionLimit < 0) with the accepted draws
// compacted across the wave.  Shading a draw costs several times its side /
// radius test, and in the per-pixel loop a wave pays for the shadow test and
// the BRDF of draw k whenever any of its 64 pixels accepts it (nearly always),
// while a quarter to a half of the lanes do.  Here each batch of up to 8 draws
// runs in three steps per wave:
//   A. every lane hashes its pixel's draws and runs the side / radius test,
//      recording the accepted slots (the seed sequence and the decisions are
//      the per-pixel loop's);
//   B. the accepted (pixel, slot) pairs, ordered pixel by pixel and draw by
//      draw, are shaded 64 at a time, one pair per lane, with the owning
//      pixel's shading inputs read from LDS: the shadow test, the reweighting
//      and the lighting, exactly as visibility() and vplLighting() compute them;
//   C. each pixel adds its own pairs' results in draw order with the per-pixel
//      loop's operations, so the outputs are bit for bit that loop's.
// Waves synchronise among their own lanes only (LDS ordered by a wavefront
// fence), so the four waves of a block run their batches independently.
constexpr int kDrawBatch = 8;

__device__ __forceinline__ void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int FMT>
__global__ __launch_bounds__(256) void gatherSamplesCompactKernel(GbufArgs g, FrameArgs fr, CubeArgs cubes, int tilesX,
                                                                 int nTiles, int samples,
                                                                 const PackedVpl* __restrict__ prep,
                                                                 const uint2* __restrict__ halfPos,
                                                                 const int32_t* __restrict__ visible,
                                                                 float* __restrict__ color,
                                                                 float* __restrict__ reservoir) {
    __shared__ Pixel pix[256];
    __shared__ uint32_t pairs[4][64 * kDrawBatch];  // (owner lane << 16) | slot
    __shared__ float4 lit[4][64];                    // a pair's lighting and its factor
    __shared__ float rew[4][64];                     // a pair's reweighting factor
    const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
    const uint32_t t = xcdSwizzle(blockIdx.x, (uint32_t)nTiles);
    const int tx = (int)t % tilesX, ty = (int)t / tilesX;
    const int x = tx * 16 + (int)threadIdx.x % 16, y = ty * 16 + (int)threadIdx.x / 16;
    const bool live = x < g.W && y < g.H;
    const size_t p = live ? (size_t)y * g.W + x : 0;
    Pixel px{};
    if (live) px = setupPixel(g, fr, x, y, p);
    pix[threadIdx.x] = px;
    f3 vplColor = splat(0.0f);
    float validSamples = 0.0f;
    const int count = min(max(visible[0], 0), SGI_MAX_TOTAL_VPLS_PER_FRAME);
    if (count > 0) {
        const float seedX = (float)x + 0.5f, seedY = (float)y + 0.5f;
        float seedZ = fr.time;
        uint32_t* wp = pairs[wave];
        for (int d0 = 0; d0 < samples; d0 += kDrawBatch) {
            const int nd = min(kDrawBatch, samples - d0);
            // A: the batch's slots hashed together and their binary16 positions fetched together
            // (eight loads in flight), then the side / radius tests; the few that the rounding
            // leaves open are decided on the FP32 record
            uint32_t slots[kDrawBatch];
            uint2 posH[kDrawBatch];
#pragma unroll
            for (int k = 0; k < kDrawBatch; ++k) {
                slots[k] = k < nd ? (uint32_t)drawSlotAll(seedX, seedY, seedZ, count) : 0u;
                posH[k] = halfPos[slots[k]];
            }
            uint32_t mask = 0, open = 0;
#pragma unroll
            for (int k = 0; k < kDrawBatch; ++k) {
                if (k < nd && live) {
                    const int dcs = sideRadiusHalf(px, posH[k]);
                    if (dcs > 0) mask |= 1u << k;
                    else if (dcs < 0) open |= 1u << k;
                }
            }
            if (open) {
#pragma unroll
                for (int k = 0; k < kDrawBatch; ++k) {
                    if (open >> k & 1u) {
                        const float4 q = prep[slots[k]].posRadius;
                        const f3 lmf = f3{q.x, q.y, q.z} - px.P;
                        const float d2 = dot(lmf, lmf);
                        if (passesSideRadius(px, lmf, d2, fsqrt(d2), dot(px.N, lmf), q.w)) mask |= 1u << k;
                    }
                }
            }

            // the wave's exclusive prefix of accepted counts orders the pairs pixel by pixel
            const int mine = __popc(mask);
            int incl = mine;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(incl, o, 64);
                if (lane >= o) incl += v;
            }
            const int start = incl - mine;
            const int total = __shfl(incl, 63, 64);
#pragma unroll
            for (int k = 0; k < kDrawBatch; ++k)
                if (mask >> k & 1u) wp[start + __popc(mask & ((1u << k) - 1u))] = ((uint32_t)lane << 16) | slots[k];
            waveSync();
            // B + C: 64 pairs at a time, the next chunk's records fetched while this one is shaded
            PackedVpl rec{};
            if (lane < total) rec = prep[wp[lane] & 0xFFFFu];
            for (int c0 = 0; c0 < total; c0 += 64) {
                const int q = c0 + lane;
                const PackedVpl here = rec;
                if (q + 64 < total) rec = prep[wp[q + 64] & 0xFFFFu];
                if (q < total) {
                    const uint32_t e = wp[q];
                    const Pixel& po = pix[wave * 64 + (int)(e >> 16)];
                    const Vpl L = unpackVpl(here);
                    float rw = 0.0f;
                    Hit h;
                    visibility<FMT>(po, fr, cubes, L, false, rw, h);
                    const f3 l = vplLighting(po, L, h.dist, h.nl);
                    lit[wave][lane] = make_float4(l.x, l.y, l.z, h.factor);
                    rew[wave][lane] = rw;
                }
                waveSync();
                // this lane's pairs within [c0, c0 + 64), in draw order
                const int lo = max(start, c0), hi = min(start + mine, c0 + 64);
                for (int r = lo; r < hi; ++r) {
                    const float4 v = lit[wave][r - c0];
                    vplColor = fmaF(f3{v.x, v.y, v.z}, v.w, vplColor);
                    validSamples += rew[wave][r - c0];
                }
                waveSync();
            }
        }
    }
    if (live) writeOut(g, p, vplColor, validSamples, color, reservoir);
}

template <int FMT>
hipError_t prepare(const CubeArgs& cubes, const sgi_vpl* vpls, const int32_t* visible, const sgi_atlas_entry* entries,
                   PackedVpl* prep, hipStream_t s) {
    hipLaunchKernelGGL(prepVplsKernel<FMT>, dim3((SGI_MAX_TOTAL_VPLS_PER_FRAME + 255) / 256), dim3(256), 0, s, vpls,
                       entries, visible, cubes, prep);
    return hipGetLastError();
}

template <int FMT>
hipError_t launchEx(const GbufArgs& g, const FrameArgs& fr, const CubeArgs& cubes, int mode, int samples,
                    int rejectionLimit, int accumulate64, int half, const sgi_vpl* vpls, const int32_t* visible,
                    const sgi_atlas_entry* entries, PackedVpl* prep, float* color, float* reservoir, hipStream_t s) {
    const int rx = (g.W + 15) / 16, ry = (g.H + 15) / 16;
    if (mode == SGI_GATHER_ALL) {
        if (accumulate64)
            hipLaunchKernelGGL((gatherAllKernel<FMT, true>), dim3(rx * ry), dim3(256), 0, s, g, fr, cubes, rx, rx * ry,
                               vpls, visible, entries, color, reservoir);
        else
            hipLaunchKernelGGL((gatherAllKernel<FMT, false>), dim3(rx * ry), dim3(256), 0, s, g, fr, cubes, rx,
                               rx * ry, vpls, visible, entries, color, reservoir);
    } else {
        const hipError_t e = prepare<FMT>(cubes, vpls, visible, entries, prep, s);
        if (e != hipSuccess) return e;
        if (half)
            hipLaunchKernelGGL((gatherSamplesKernel<FMT, true>), dim3(rx * ry), dim3(256), 0, s, g, fr, cubes, rx,
                               rx * ry, samples, rejectionLimit, prep, visible, color, reservoir);
        else if (rejectionLimit < 0)
            hipLaunchKernelGGL((gatherSamplesCompactKernel<FMT>), dim3(rx * ry), dim3(256), 0, s, g, fr, cubes, rx,
                               rx * ry, samples, prep, halfPositions(prep), visible, color, reservoir);
        else
            hipLaunchKernelGGL((gatherSamplesKernel<FMT, false>), dim3(rx * ry), dim3(256), 0, s, g, fr, cubes, rx,
                               rx * ry, samples, rejectionLimit, prep, visible, color, reservoir);
    }
    return hipGetLastError();
}

template <int FMT, bool HALF>
hipError_t launchFmt(const GbufArgs& g, const FrameArgs& fr, const CubeArgs& cubes, int tileSize, int tilesX,
                     int tilesY, int K, const sgi_vpl* vpls, const int32_t* visible, const sgi_atlas_entry* entries,
                     const int32_t* tiles, PackedVpl* prep, float* color, float* reservoir, hipStream_t s) {
    if (!tiles) {
        const int rx = (g.W + 15) / 16, ry = (g.H + 15) / 16;
        const hipError_t e = prepare<FMT>(cubes, vpls, visible, entries, prep, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((gatherRandomKernel<FMT, HALF>), dim3(rx * ry), dim3(256), 0, s, g, fr, cubes, rx, rx * ry,
                           prep, visible, color, reservoir);
        return hipGetLastError();
    }
    const int nTiles = tilesX * tilesY;
    switch (tileSize) {
        case 8:
            hipLaunchKernelGGL((gatherTilesKernel<FMT, 8, HALF>), dim3(nTiles), dim3(64), 0, s, g, fr, cubes, tilesX,
                               nTiles, K, vpls, entries, tiles, color, reservoir);
            break;
        case 16:
            hipLaunchKernelGGL((gatherTilesKernel<FMT, 16, HALF>), dim3(nTiles), dim3(256), 0, s, g, fr, cubes, tilesX,
                               nTiles, K, vpls, entries, tiles, color, reservoir);
            break;
        case 32:
            hipLaunchKernelGGL((gatherTilesKernel<FMT, 32, HALF>), dim3(nTiles), dim3(1024), 0, s, g, fr, cubes,
                               tilesX, nTiles, K, vpls, entries, tiles, color, reservoir);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int FMT>
hipError_t launchPrecision(const GbufArgs& g, const FrameArgs& fr, const CubeArgs& cubes, int tileSize, int tilesX,
                           int tilesY, int K, const sgi_vpl* vpls, const int32_t* visible,
                           const sgi_atlas_entry* entries, const int32_t* tiles, int precision, PackedVpl* prep,
                           float* color, float* reservoir, hipStream_t s) {
    if (precision == SGI_PRECISION_FP16)
        return launchFmt<FMT, true>(g, fr, cubes, tileSize, tilesX, tilesY, K, vpls, visible, entries, tiles, prep,
                                    color, reservoir, s);
    return launchFmt<FMT, false>(g, fr, cubes, tileSize, tilesX, tilesY, K, vpls, visible, entries, tiles, prep,
                                 color, reservoir, s);
}

}  // namespace

namespace sgi_launch {

size_t packedVplBytes() { return sizeof(PackedVpl) + sizeof(uint2); }  // a record and its binary16 position

hipError_t gather(const GbufArgs& g, const FrameArgs& fr, const CubeArgs& cubes, int format, int tileSize, int tilesX,
                  int tilesY, int K, const sgi_vpl* vpls, const int32_t* visible, const sgi_atlas_entry* entries,
                  const int32_t* tiles, int precision, void* prep, float* color, float* reservoir, hipStream_t s) {
    if (precision != SGI_PRECISION_FP32 && precision != SGI_PRECISION_FP16) return hipErrorNotSupported;
    PackedVpl* pv = static_cast<PackedVpl*>(prep);
    if (format == SGI_TEXEL_F16)
        return launchPrecision<SGI_TEXEL_F16>(g, fr, cubes, tileSize, tilesX, tilesY, K, vpls, visible, entries,
                                              tiles, precision, pv, color, reservoir, s);
    if (format == SGI_TEXEL_UNORM16)
        return launchPrecision<SGI_TEXEL_UNORM16>(g, fr, cubes, tileSize, tilesX, tilesY, K, vpls, visible, entries,
                                                  tiles, precision, pv, color, reservoir, s);
    return launchPrecision<SGI_TEXEL_F32>(g, fr, cubes, tileSize, tilesX, tilesY, K, vpls, visible, entries, tiles,
                                          precision, pv, color, reservoir, s);
}

hipError_t gatherEx(const GbufArgs& g, const FrameArgs& fr, const CubeArgs& cubes, int format, int mode, int samples,
                    int rejectionLimit, int accumulate64, int half, const sgi_vpl* vpls, const int32_t* visible,
                    const sgi_atlas_entry* entries, void* prep, float* color, float* reservoir, hipStream_t s) {
    PackedVpl* pv = static_cast<PackedVpl*>(prep);
    if (format == SGI_TEXEL_F16)
        return launchEx<SGI_TEXEL_F16>(g, fr, cubes, mode, samples, rejectionLimit, accumulate64, half, vpls, visible,
                                       entries, pv, color, reservoir, s);
    if (format == SGI_TEXEL_UNORM16)
        return launchEx<SGI_TEXEL_UNORM16>(g, fr, cubes, mode, samples, rejectionLimit, accumulate64, half, vpls,
                                           visible, entries, pv, color, reservoir, s);
    return launchEx<SGI_TEXEL_F32>(g, fr, cubes, mode, samples, rejectionLimit, accumulate64, half, vpls, visible,
                                   entries, pv, color, reservoir, s);
}

hipError_t gatherViews(const sgi_batch::ViewDesc* views, int numViews, int totalTiles, int tileSize, int K,
                       const CubeArgs& cubes, int format, hipStream_t s) {
    const int threads = tileSize * tileSize;
#define SGI_GATHER_VIEWS(F, T)                                                                                      \
    hipLaunchKernelGGL((gatherTilesViewsKernel<F, T>), dim3(totalTiles), dim3(threads), 0, s, views, numViews,     \
                       totalTiles, K, cubes)
    if (format == SGI_TEXEL_F16) {
        if (tileSize == 8) SGI_GATHER_VIEWS(SGI_TEXEL_F16, 8);
        else if (tileSize == 16) SGI_GATHER_VIEWS(SGI_TEXEL_F16, 16);
        else if (tileSize == 32) SGI_GATHER_VIEWS(SGI_TEXEL_F16, 32);
        else return hipErrorInvalidValue;
    } else if (format == SGI_TEXEL_UNORM16) {
        if (tileSize == 8) SGI_GATHER_VIEWS(SGI_TEXEL_UNORM16, 8);
        else if (tileSize == 16) SGI_GATHER_VIEWS(SGI_TEXEL_UNORM16, 16);
        else if (tileSize == 32) SGI_GATHER_VIEWS(SGI_TEXEL_UNORM16, 32);
        else return hipErrorInvalidValue;
    } else {
        if (tileSize == 8) SGI_GATHER_VIEWS(SGI_TEXEL_F32, 8);
        else if (tileSize == 16) SGI_GATHER_VIEWS(SGI_TEXEL_F32, 16);
        else if (tileSize == 32) SGI_GATHER_VIEWS(SGI_TEXEL_F32, 32);
        else return hipErrorInvalidValue;
    }
#undef SGI_GATHER_VIEWS
    return hipGetLastError();
}

}  // namespace sgi_launch
