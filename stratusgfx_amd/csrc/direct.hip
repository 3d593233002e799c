// Deferred direct sun lighting — pbr.fs:82-145 with INFINITE_LIGHTING_ENABLED
// and no point lights: 3x3 PCF over two blended cascades
// (calculateInfiniteShadowValue, pbr.glsl:158-227) times BRDF_Burley
// (pbr2.glsl:194-250) through calculateLighting_Burley (pbr2.glsl:454-478).
// One thread per pixel; blocks cover 16x16 tiles, XCD-contiguous.
#include "launch.h"
#include "views.h"

namespace {

using namespace sgi_dev;

__device__ __forceinline__ void directBody(const GbufArgs& g, const FrameArgs& fr, const CsmArgs& csm, int t,
                                           int tilesX, float* __restrict__ lit) {
    const int x = (t % tilesX) * 16 + (int)threadIdx.x % 16, y = (t / tilesX) * 16 + (int)threadIdx.x / 16;
    if (x >= g.W || y >= g.H) return;
    const size_t p = (size_t)y * g.W + x;
    const f3 fragPos = worldPositionFromDepth(texelCentre(x, g.W), texelCentre(y, g.H), g.depth[p], fr.invPV);
    const f3 viewMinusFrag = ld3(fr.viewPos) - fragPos;
    const f3 V = normalize(viewMinusFrag);
    const float viewDist = length(viewMinusFrag);
    const float4 al = gAlbedo(g, p);
    const f3 baseColor = f3{al.x, al.y, al.z};
    const f3 N = gNormal(g, p);
    const f3 rma = gRma(g, p);
    const float roughness = rma.x, metallic = rma.y;
    const float ambient = g.ssao ? g.ssao[p] : 1.0f;
    const float2 refl = gRefl(g, p);
    const float reflR = refl.x, reflG = refl.y;
    const f3 emissive = f3{al.w, reflG, rma.z};

    const f3 sun = ld3(fr.sunDir);
    const f3 blends = cascadeBlends(csm, fragPos);
    const float shadowFactor = infiniteShadowValueBlock(csm, sun, fragPos, blends, N, csm.bias);
    // calculateDirectionalLighting passes 1 - shadowFactor, which the Burley path inverts again
    const float sf = 1.0f - (1.0f - shadowFactor);

    // BRDF_Burley with diffuseDivisor = specularMultiplier = 1
    const float a = remapRoughness(roughness, fr);
    const f3 diffuseColor = ((1.0f - metallic) * baseColor) / (splat(1.0f) + kEPS);
    const f3 baseRefl = splat(reflR);
    const f3 f0 = (0.16f * baseRefl * baseRefl * (1.0f - metallic) + baseColor * metallic) * splat(1.0f);
    const f3 L = normalize(sun);
    const f3 H = normalize(V + L);
    const float NdotV = saturate(dot(N, V)), NdotL = saturate(dot(N, L));
    const float NdotH = saturate(dot(N, H)), LdotH = saturate(dot(L, H));
    const f3 Fr = (ndfGGX(NdotH, a) * visSmithGGX(NdotV, NdotL, a)) * fresnelSchlick(LdotH, f0);
    // Fd_Burley (pbr2.glsl:91-96)
    const float f90 = 0.5f + 2.0f * a * LdotH * LdotH;
    const float lightScatter = 1.0f + (f90 - 1.0f) * pow5(1.0f - NdotL);
    const float viewScatter = 1.0f + (f90 - 1.0f) * pow5(1.0f - NdotV);
    const f3 Fd = diffuseColor * (lightScatter * viewScatter * (1.0f / kPI));
    const f3 brdf = (Fd + Fr) * NdotL;

    const f3 sunColor = ld3(fr.sunColor);
    const f3 ambientTerm = brdf * ambient * sunColor * 0.0f;  // ambientIntensity 0
    const f3 finalBrightness = brdf * sunColor;
    const float f = expf(-fr.fogDensity * viewDist);  // applyFog with fog intensity 0
    const f3 fogged = finalBrightness * f + (ld3(fr.fogColor) * 0.0f) * (1.0f - f);
    f3 color = splat(0.0f);
    color = color + 1.0f * (ambientTerm + sf * fogged);
    const f3 outc = color + emissive * fr.emissionStrength;
    lit[3 * p + 0] = gmin(outc.x, 65504.0f);
    lit[3 * p + 1] = gmin(outc.y, 65504.0f);
    lit[3 * p + 2] = gmin(outc.z, 65504.0f);
}

__global__ __launch_bounds__(256) void directKernel(GbufArgs g, FrameArgs fr, CsmArgs csm, int tilesX, int nTiles,
                                                    float* __restrict__ lit) {
    directBody(g, fr, csm, (int)xcdSwizzle(blockIdx.x, (uint32_t)nTiles), tilesX, lit);
}

// Batched: the 16x16 tiles of every view, numbered view after view.
__global__ __launch_bounds__(256) void directViewsKernel(const sgi_batch::ViewDesc* __restrict__ views, int numViews,
                                                         int total) {
    const int gt = (int)xcdSwizzle(blockIdx.x, (uint32_t)total);
    const sgi_batch::ViewDesc& d = views[sgi_batch::findView<sgi_batch::kDirectTiles>(views, numViews, gt)];
    directBody(d.g, d.fr, d.csm, gt - d.dtileStart, (d.g.W + 15) / 16, d.direct);
}

}  // namespace

namespace sgi_launch {

hipError_t directLight(const GbufArgs& g, const FrameArgs& fr, const CsmArgs& csm, float* lit, hipStream_t s) {
    const int tx = (g.W + 15) / 16, ty = (g.H + 15) / 16;
    hipLaunchKernelGGL(directKernel, dim3(tx * ty), dim3(256), 0, s, g, fr, csm, tx, tx * ty, lit);
    return hipGetLastError();
}

hipError_t directLightViews(const sgi_batch::ViewDesc* views, int numViews, int totalTiles16, hipStream_t s) {
    hipLaunchKernelGGL(directViewsKernel, dim3(totalTiles16), dim3(256), 0, s, views, numViews, totalTiles16);
    return hipGetLastError();
}

}  // namespace sgi_launch
