// Deferred direct sun lighting — pbr.fs:82-145 with INFINITE_LIGHTING_ENABLED
// and no point lights: 3x3 PCF over two blended cascades
// (calculateInfiniteShadowValue, pbr.glsl:158-227) times BRDF_Burley
// (pbr2.glsl:194-250) through calculateLighting_Burley (pbr2.glsl:454-478).
// One thread per pixel; blocks cover 16x16 tiles, XCD-contiguous.  Not built
// with the exact-contraction rule of the other passes where it says so below:
// only the shading after the shadow factor is contracted.
#include "launch.h"
#include "views.h"

namespace {

using namespace sgi_dev;

// The values the reference decides with are exact (the fragment position,
// the decoded normal and the cascade test they feed); the continuous shading
// after the shadow factor uses the hardware reciprocal, square root and exp
// with FMA contraction, within a few ulps of the exact restatement.
__device__ __forceinline__ void directBody(const GbufArgs& g, const FrameArgs& fr, const CsmArgs& csm, int t,
                                           int tilesX, float* __restrict__ lit) {
    const int x = (t % tilesX) * 16 + (int)threadIdx.x % 16, y = (t / tilesX) * 16 + (int)threadIdx.x / 16;
    if (x >= g.W || y >= g.H) return;
    const size_t p = (size_t)y * g.W + x;
    const f3 fragPos = worldPositionFromDepth(texelCentre(x, g.W), texelCentre(y, g.H), g.depth[p], fr.invPV);
    const f3 N = gNormal(g, p);
    const f3 sun = ld3(fr.sunDir);
    const f3 blends = cascadeBlends(csm, fragPos);
    const float shadowFactor = infiniteShadowValueBlock(csm, sun, fragPos, blends, N, csm.bias);
    // calculateDirectionalLighting passes 1 - shadowFactor, which the Burley path inverts again
    const float sf = 1.0f - (1.0f - shadowFactor);

    const f3 viewMinusFrag = ld3(fr.viewPos) - fragPos;
    const float v2 = dotF(viewMinusFrag, viewMinusFrag);
    const float vInv = frsq(v2);
    const f3 V = viewMinusFrag * vInv;
    const float viewDist = v2 * vInv;
    const float4 al = gAlbedo(g, p);
    const f3 baseColor = f3{al.x, al.y, al.z};
    const f3 rma = gRma(g, p);
    const float roughness = rma.x, metallic = rma.y;
    const float ambient = g.ssao ? g.ssao[p] : 1.0f;
    const float2 refl = gRefl(g, p);
    const float reflR = refl.x, reflG = refl.y;
    const f3 emissive = f3{al.w, reflG, rma.z};

    // BRDF_Burley (pbr2.glsl:194-250) with diffuseDivisor = specularMultiplier = 1
    const float a = remapRoughness(roughness, fr);
    const f3 diffuseColor = ((1.0f - metallic) * baseColor) * frcp(1.0f + kEPS);
    const float f0s = 0.16f * reflR * reflR * (1.0f - metallic);
    const f3 f0 = fmaF(baseColor, metallic, splat(f0s));
    const f3 L = sun * frsq(dotF(sun, sun));
    const f3 VL = V + L;
    const f3 H = VL * frsq(dotF(VL, VL));
    const float NdotV = saturate(dotF(N, V)), NdotL = saturate(dotF(N, L));
    const float NdotH = saturate(dotF(N, H)), LdotH = saturate(dotF(L, H));
    // NDF_GGX and Visibility_G_SmithGGX (pbr2.glsl:51-64), Fr_Fresnel_Schlick (:73-76)
    const float r2 = a * a;
    const float fh = __builtin_fmaf(__builtin_fmaf(NdotH, r2, -NdotH), NdotH, 1.0f);
    const float ndf = r2 * frcp(gmax(kPI * fh * fh, kEPS));
    const float ggxv = NdotL * fsqrt(__builtin_fmaf(__builtin_fmaf(-r2, NdotV, NdotV), NdotV, r2));
    const float ggxl = NdotV * fsqrt(__builtin_fmaf(__builtin_fmaf(-r2, NdotL, NdotL), NdotL, r2));
    const float vis = 0.5f * frcp(gmax(ggxv + ggxl, kEPS));
    const float fr5 = pow5(1.0f - LdotH);
    const f3 Fr = fmaF(f0, 1.0f - fr5, splat(fr5)) * (ndf * vis);
    // Fd_Burley (pbr2.glsl:91-96)
    const float f90 = __builtin_fmaf(2.0f * a * LdotH, LdotH, 0.5f);
    const float lightScatter = __builtin_fmaf(f90 - 1.0f, pow5(1.0f - NdotL), 1.0f);
    const float viewScatter = __builtin_fmaf(f90 - 1.0f, pow5(1.0f - NdotV), 1.0f);
    const f3 Fd = diffuseColor * (lightScatter * viewScatter * (1.0f / kPI));
    const f3 brdf = (Fd + Fr) * NdotL;

    const f3 sunColor = ld3(fr.sunColor);
    const f3 ambientTerm = brdf * ambient * sunColor * 0.0f;  // ambientIntensity 0
    const f3 finalBrightness = brdf * sunColor;
    const float f = fexp(-fr.fogDensity * viewDist);  // applyFog with fog intensity 0
    const f3 fogged = finalBrightness * f + (ld3(fr.fogColor) * 0.0f) * (1.0f - f);
    const f3 outc = (ambientTerm + sf * fogged) + emissive * fr.emissionStrength;
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
