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

namespace {

using namespace sgi_dev;

struct Pixel {
    f3 P, V, N, f0, diffColor;
    float distToCamera, a, NdotV, ao, fog;
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
    o.cube = cubeLayerBase(cubes, en.index, en.layer, FMT == SGI_TEXEL_F16 ? 2 : 4);
    return o;
}

// vplDiffuseAttenuation (pbr.glsl:260-267) with the light distance known.
__device__ __forceinline__ float vplAttenuationFast(float lightDist, float radius, float invRadius) {
    const float ratio = gmin(lightDist * invRadius, 1.0f);
    const float minDist = mixv(10.0f, 1.0f, fexp(-3.0f * ratio)) * radius;
    return frcp(minDist + 1.0f * lightDist * lightDist);
}

// The continuous part of one VPL's lighting: calculateVirtualPointLighting2 ->
// calculateLighting_Lambert_SeparateDiffuseSpecular (pbr2.glsl:541-574, 613-657),
// with FMA contraction allowed (no decision depends on these values).
__device__ __forceinline__ f3 vplLighting(const Pixel& px, const Vpl& L, f3 lmf, float dist) {
#pragma clang fp contract(fast)
    const f3 specDir = L.spec - px.P;
    const float s2 = dot(specDir, specDir);
    const float sInv = frsq(s2);
    const float sLen = s2 * sInv;
    const f3 Ls = specDir * sInv;
    const f3 h = px.V + Ls;
    const f3 Hs = h * frsq(dot(h, h));
    const float NdotL = saturate(dot(px.N, Ls));
    const float NdotH = saturate(dot(px.N, Hs));
    const float LdotH = saturate(dot(Ls, Hs));
    const float r2 = px.a * px.a;  // NDF_GGX / Visibility_G_SmithGGX (pbr2.glsl:51-64)
    const float f = (NdotH * r2 - NdotH) * NdotH + 1.0f;
    const float ndf = r2 * frcp(gmax(kPI * f * f, kEPS));
    const float ggxv = NdotL * fsqrt((px.NdotV - r2 * px.NdotV) * px.NdotV + r2);
    const float ggxl = px.NdotV * fsqrt((NdotL - r2 * NdotL) * NdotL + r2);
    const float vis = 0.5f * frcp(gmax(ggxv + ggxl, kEPS));
    const f3 Fr = (ndf * vis) * fresnelSchlick(LdotH, px.f0) * NdotL;
    const float invDist = frcp(dist);
    const float NdotLd = saturate(dot(px.N, lmf) * invDist);
    const f3 Fd = px.diffColor * (1.0f / kPI) * NdotLd;
    const f3 brdf = vplAttenuationFast(sLen, L.radius, L.invRadius) * Fr +
                    vplAttenuationFast(dist, L.radius, L.invRadius) * Fd;
    // the ambient term is brdf * ao * colour * 0 and the fog intensity is 0, so
    // the lighting is (brdf * colour) * exp(-fogDensity * viewDist)
    return (brdf * L.color) * px.fog;
}

// One VPL's contribution (the body of the sample loop, vpl_pbr_gi.fs:149-199).
// Returns false when the side / radius test rejects it.
template <int FMT>
__device__ __forceinline__ bool evalVpl(const Pixel& px, const FrameArgs& fr, const CubeArgs& cubes, const Vpl& L,
                                        bool allowReject, f3& vplColor, float& validSamples) {
    const f3 lmf = L.pos - px.P;
    const float dist = length(lmf);  // exact: radius test and shadow compare
    if (allowReject && (normalizedDotNegative(px.N, lmf, dist) || dist > L.radius)) return false;
    const float distanceRatio = clampv((2.0f * dist) * L.invRadius, 0.0f, 1.0f);
    float shadowFactor = 0.0f;
    if (px.distToCamera < 500.0f) {  // calculateShadowValue1Sample, pbr.glsl:116-136
        const f3 fragDir = px.P - L.pos;
        const float currentDepth = dist;  // |P - L| == |L - P| bit for bit
        const float bias = currentDepth * gmax(0.1f, saturate(dot(lmf, px.N)) * 0.03f);
        const float depth = cubeFetch<FMT>(L.cube, cubes.res, fragDir) * L.farPlane;
        shadowFactor = (currentDepth - bias) > depth ? 1.0f : 0.0f;
    }
    // min(shadow, mix(minGiOcc, 1, ratio)) > 0 exactly when the shadow test fired
    shadowFactor = gmin(shadowFactor, mixv(fr.minGiOcc, 1.0f, distanceRatio));
    float reweightingFactor = 1.0f;
    if (shadowFactor > 0.0f) reweightingFactor = (1.0f - distanceRatio) * fr.minGiOcc + distanceRatio;
    validSamples += reweightingFactor;
    vplColor = vplColor + (1.0f - shadowFactor) * (px.ao * vplLighting(px, L, lmf, dist));
    return true;
}

__device__ __forceinline__ Pixel setupPixel(const GbufArgs& g, const FrameArgs& fr, int x, int y, size_t p) {
    Pixel px;
    px.P = worldPositionFromDepth(texelCentre(x, g.W), texelCentre(y, g.H), g.depth[p], fr.invPV);
    const f3 vp = ld3(fr.viewPos) - px.P;
    const float vp2 = dot(vp, vp);
    px.distToCamera = __builtin_sqrtf(vp2);  // exact: shadow on / off and the sample count
    px.V = vp * frsq(vp2);
    const float* al = g.albedo + 4 * p;
    const f3 baseColor = f3{gmax(al[0], 0.01f), gmax(al[1], 0.01f), gmax(al[2], 0.01f)};
    px.N = decodeNormal(g.normal + 3 * p);  // exact: side test
    const float* rma = g.rma + 3 * p;
    const float roughness = gmax(0.5f, rma[0]);
    const float metallic = rma[1];
    px.ao = clampv(g.ssao ? g.ssao[p] : 1.0f, 0.35f, 1.0f);
    const f3 baseRefl = splat(g.refl[2 * p]);
    // BRDF_Base with diffuseDivisor = baseColor, specularMultiplier = 1 / (baseColor + eps)
    px.a = remapRoughness(roughness, fr);
    const f3 specMul = f3{frcp(baseColor.x + kEPS), frcp(baseColor.y + kEPS), frcp(baseColor.z + kEPS)};
    px.diffColor = ((1.0f - metallic) * baseColor) * specMul;
    px.f0 = (0.16f * baseRefl * baseRefl * (1.0f - metallic) + baseColor * metallic) * specMul;
    px.NdotV = saturate(dot(px.N, px.V));
    px.fog = fexp(-fr.fogDensity * px.distToCamera);
    return px;
}

__device__ __forceinline__ void writeOut(const GbufArgs& g, size_t p, f3 vplColor, float validSamples, float* color,
                                         float* reservoir) {
    const float* al = g.albedo + 4 * p;
    color[3 * p + 0] = gmax(al[0], 0.01f) + kEPS;
    color[3 * p + 1] = gmax(al[1], 0.01f) + kEPS;
    color[3 * p + 2] = gmax(al[2], 0.01f) + kEPS;
    float4 r;
    r.x = gmin(vplColor.x, 65504.0f);
    r.y = gmin(vplColor.y, 65504.0f);
    r.z = gmin(vplColor.z, 65504.0f);
    r.w = gmax(validSamples, 1.0f);
    reinterpret_cast<float4*>(reservoir)[p] = r;
}

constexpr int kMaxK = SGI_MAX_VPLS_PER_TILE_LIMIT;

template <int FMT, int TS>
__global__ __launch_bounds__(TS * TS) void gatherTilesKernel(GbufArgs g, FrameArgs fr, CubeArgs cubes, int tilesX,
                                                            int nTiles, int K, const sgi_vpl* __restrict__ vpls,
                                                            const sgi_atlas_entry* __restrict__ entries,
                                                            const int32_t* __restrict__ tiles,
                                                            float* __restrict__ color, float* __restrict__ reservoir) {
    __shared__ Vpl lv[kMaxK];
    __shared__ int ln;
    const uint32_t t = xcdSwizzle(blockIdx.x, (uint32_t)nTiles);
    const int tx = (int)t % tilesX, ty = (int)t / tilesX;
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
    for (int k = 0; k < n; ++k) evalVpl<FMT>(px, fr, cubes, lv[k], true, vplColor, validSamples);
    writeOut(g, p, vplColor, validSamples, color, reservoir);
}

template <int FMT>
__global__ __launch_bounds__(256) void gatherRandomKernel(GbufArgs g, FrameArgs fr, CubeArgs cubes, int tilesX,
                                                         int nTiles, const sgi_vpl* __restrict__ vpls,
                                                         const int32_t* __restrict__ visible,
                                                         const sgi_atlas_entry* __restrict__ entries,
                                                         float* __restrict__ color, float* __restrict__ reservoir) {
    const uint32_t t = xcdSwizzle(blockIdx.x, (uint32_t)nTiles);
    const int tx = (int)t % tilesX, ty = (int)t / tilesX;
    const int x = tx * 16 + (int)threadIdx.x % 16, y = ty * 16 + (int)threadIdx.x / 16;
    if (x >= g.W || y >= g.H) return;
    const size_t p = (size_t)y * g.W + x;
    const Pixel px = setupPixel(g, fr, x, y, p);
    f3 vplColor = splat(0.0f);
    float validSamples = 0.0f;
    const int numVisible = visible[0];
    if (numVisible > 0) {
        const float* rma = g.rma + 3 * p;
        const float roughness = gmax(0.5f, rma[0]);
        const float roughnessWeight = 1.0f - roughness;
        const float distRatioToCamera = gmin(1.0f - px.distToCamera / 1000.0f, 1.0f);
        const int maxSamplesPerPixel = (int)mixv(2.0f, 10.0f, roughnessWeight);
        int samplesMax = (int)((float)maxSamplesPerPixel * distRatioToCamera);
        samplesMax = samplesMax > 1 ? samplesMax : 1;
        const int maxRandomIndex = numVisible - 1;
        const float seedX = (float)x + 0.5f, seedY = (float)y + 0.5f;
        float seedZ = fr.time;
        for (int i = 0, resamples = 0; i < samplesMax; i += 1) {
            const float rnd = random3(seedX, seedY, seedZ);
            seedZ += 10000.0f;
            const int lightIndex = (int)((float)maxRandomIndex * rnd);
            const Vpl L = loadVpl<FMT>(vpls + lightIndex, entries + lightIndex, cubes);
            if (!evalVpl<FMT>(px, fr, cubes, L, resamples < 5, vplColor, validSamples)) {
                ++resamples;
                --i;
                continue;
            }
        }
    }
    writeOut(g, p, vplColor, validSamples, color, reservoir);
}

template <int FMT>
hipError_t launchFmt(const GbufArgs& g, const FrameArgs& fr, const CubeArgs& cubes, int tileSize, int tilesX,
                     int tilesY, int K, const sgi_vpl* vpls, const int32_t* visible, const sgi_atlas_entry* entries,
                     const int32_t* tiles, float* color, float* reservoir, hipStream_t s) {
    if (!tiles) {
        const int rx = (g.W + 15) / 16, ry = (g.H + 15) / 16;
        hipLaunchKernelGGL(gatherRandomKernel<FMT>, dim3(rx * ry), dim3(256), 0, s, g, fr, cubes, rx, rx * ry, vpls,
                           visible, entries, color, reservoir);
        return hipGetLastError();
    }
    const int nTiles = tilesX * tilesY;
    switch (tileSize) {
        case 8:
            hipLaunchKernelGGL((gatherTilesKernel<FMT, 8>), dim3(nTiles), dim3(64), 0, s, g, fr, cubes, tilesX, nTiles,
                               K, vpls, entries, tiles, color, reservoir);
            break;
        case 16:
            hipLaunchKernelGGL((gatherTilesKernel<FMT, 16>), dim3(nTiles), dim3(256), 0, s, g, fr, cubes, tilesX,
                               nTiles, K, vpls, entries, tiles, color, reservoir);
            break;
        case 32:
            hipLaunchKernelGGL((gatherTilesKernel<FMT, 32>), dim3(nTiles), dim3(1024), 0, s, g, fr, cubes, tilesX,
                               nTiles, K, vpls, entries, tiles, color, reservoir);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

namespace sgi_launch {

hipError_t gather(const GbufArgs& g, const FrameArgs& fr, const CubeArgs& cubes, int format, int tileSize, int tilesX,
                  int tilesY, int K, const sgi_vpl* vpls, const int32_t* visible, const sgi_atlas_entry* entries,
                  const int32_t* tiles, int precision, float* color, float* reservoir, hipStream_t s) {
    if (precision != SGI_PRECISION_FP32) return hipErrorNotSupported;
    if (format == SGI_TEXEL_F16)
        return launchFmt<SGI_TEXEL_F16>(g, fr, cubes, tileSize, tilesX, tilesY, K, vpls, visible, entries, tiles,
                                        color, reservoir, s);
    return launchFmt<SGI_TEXEL_F32>(g, fr, cubes, tileSize, tilesX, tilesY, K, vpls, visible, entries, tiles, color,
                                    reservoir, s);
}

}  // namespace sgi_launch
