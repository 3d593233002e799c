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

using namespace sgi_dev;

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

// ---- TAA history (PerformTaaPostFx_'s copy, StratusRendererBackend.cpp:2342-2349)
//
// The TAA target is RGBA16F (:708-711): the output rounds to binary16 (RNE);
// the NEAREST blit into the RGB8 UNORM previousFrameBuffer (:449-453) clamps
// to [0, 1] and stores floor(v * 255 + 0.5); the next frame reads k / 255.
__global__ __launch_bounds__(256) void taaHistoryKernel(const float* __restrict__ taa, long n3,
                                                         float* __restrict__ prev) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n3) return;
    const float h = __half2float(__float2half_rn(taa[q]));
    const uint32_t k = (uint32_t)(sat(h) * 255.0f + 0.5f);
    prev[q] = (float)k / 255.0f;
}

// ---- FXAA (PerformFxaaPostFx_, StratusRendererBackend.cpp:2293-2321)
//
// The reference runs two passes: fxaa_luminance.fs writes the tonemapped
// colour with its luminance in alpha to an RGBA8 target, fxaa_smoothing.fs
// filters that target.  Here one kernel does both: every tap of the smoothing
// pass recomputes the stored alpha code of the texels it reads from their
// tonemapped rgb codes (the luminance in the shader's order, then the UNORM8
// store), which is exactly what the intermediate target would hold, so the
// intermediate never touches HBM (4 B read and 4 B written per pixel; the
// neighbourhood and edge-walk taps hit L2).  Built with -ffp-contract=off, so
// every product and sum rounds as the shaders write them.
struct FxaaImage {
    const uint32_t* px;  // RGBA8 texels [H][W] of the tonemapped target
    int W, H;
};

__device__ __forceinline__ uint32_t fxaaTexel(const FxaaImage& im, int x, int y) {
    return im.px[(size_t)clampi(y, 0, im.H - 1) * im.W + clampi(x, 0, im.W - 1)];
}

// k / 255 correctly rounded, from a 256-entry table in LDS (filled per block).
__device__ __forceinline__ float unorm8(const float* lut, uint32_t texel, int ch) {
    return lut[(texel >> (8 * ch)) & 0xFFu];
}

__device__ __forceinline__ uint32_t store8(float v) { return (uint32_t)(sat(v) * 255.0f + 0.5f); }

// The alpha of the luminance target at texel (x, y): aa_common.glsl:7-9 on the
// saturated colour (already in [0, 1]), stored UNORM8, read back as k / 255.
__device__ __forceinline__ float lumaAt(const FxaaImage& im, const float* lut, int x, int y) {
    const uint32_t t = fxaaTexel(im, x, y);
    const float lum = 0.2126f * unorm8(lut, t, 0) + 0.7152f * unorm8(lut, t, 1) + 0.0722f * unorm8(lut, t, 2);
    return lut[store8(lum)];
}

// GL LINEAR, CLAMP_TO_EDGE, of the luminance target's alpha at (u, v).
__device__ __forceinline__ float lumaLinear(const FxaaImage& im, const float* lut, float u, float v) {
    const float fu = u * (float)im.W - 0.5f, fv = v * (float)im.H - 0.5f;
    const float i0f = floorf(fu), j0f = floorf(fv);
    const float a = fu - i0f, b = fv - j0f;
    const int ia = texelIndex(i0f, im.W), ib = texelIndex(i0f + 1.0f, im.W);
    const int ja = texelIndex(j0f, im.H), jb = texelIndex(j0f + 1.0f, im.H);
    const float t00 = lumaAt(im, lut, ia, ja), t10 = lumaAt(im, lut, ib, ja);
    const float t01 = lumaAt(im, lut, ia, jb), t11 = lumaAt(im, lut, ib, jb);
    return (1.0f - a) * (1.0f - b) * t00 + a * (1.0f - b) * t10 + (1.0f - a) * b * t01 + a * b * t11;
}

__global__ __launch_bounds__(256) void fxaaKernel(FxaaImage im, float contrastThreshold, float relativeThreshold,
                                                   float subpixelBlending, uint32_t* __restrict__ out) {
    __shared__ float lut[256];
    lut[threadIdx.x] = (float)threadIdx.x / 255.0f;
    __syncthreads();
    const int x = blockIdx.x * 16 + (int)threadIdx.x % 16, y = blockIdx.y * 16 + (int)threadIdx.x / 16;
    if (x >= im.W || y >= im.H) return;
    const size_t p = (size_t)y * im.W + x;
    const uint32_t centre = im.px[p];
    const float u = texelCentre(x, im.W), v = texelCentre(y, im.H);
    const float lumaCenter = lumaAt(im, lut, x, y);
    const float lumaRight = lumaAt(im, lut, x + 1, y), lumaLeft = lumaAt(im, lut, x - 1, y);
    const float lumaTop = lumaAt(im, lut, x, y + 1), lumaBot = lumaAt(im, lut, x, y - 1);
    const float lumaMax = fmaxf(fmaxf(fmaxf(fmaxf(lumaCenter, lumaRight), lumaLeft), lumaTop), lumaBot);
    const float lumaMin = fminf(fminf(fminf(fminf(lumaCenter, lumaRight), lumaLeft), lumaTop), lumaBot);
    const float contrast = lumaMax - lumaMin;
    const float threshold = fmaxf(contrastThreshold, relativeThreshold * lumaMax);
    if (contrast < threshold) {
        // the luminance pass re-stores k / 255 as k: the colour is the input's
        out[p] = (centre & 0x00FFFFFFu) | 0xFF000000u;
        return;
    }
    const float lumaTopRight = lumaAt(im, lut, x + 1, y + 1), lumaBotRight = lumaAt(im, lut, x + 1, y - 1);
    const float lumaTopLeft = lumaAt(im, lut, x - 1, y + 1), lumaBotLeft = lumaAt(im, lut, x - 1, y - 1);
    float average = 2.0f * (lumaRight + lumaLeft + lumaTop + lumaBot) + lumaTopRight + lumaBotRight + lumaTopLeft +
                    lumaBotLeft;
    average = average / 12.0f;
    const float centerAvgContrast = fabsf(average - lumaCenter);
    const float t = sat(sat(centerAvgContrast / contrast));
    float pixelBlendFactor = t * t * (3.0f - 2.0f * t);  // smoothstep(0, 1, .)
    pixelBlendFactor = pixelBlendFactor * pixelBlendFactor * subpixelBlending;
    const float horizontal = fabsf(lumaTop + lumaBot - 2.0f * lumaCenter) * 2.0f +
                             fabsf(lumaTopRight + lumaBotRight - 2.0f * lumaRight) +
                             fabsf(lumaTopLeft + lumaBotLeft - 2.0f * lumaLeft);
    const float vertical = fabsf(lumaRight + lumaLeft - 2.0f * lumaCenter) * 2.0f +
                           fabsf(lumaTopRight + lumaTopLeft - 2.0f * lumaTop) +
                           fabsf(lumaBotRight + lumaBotLeft - 2.0f * lumaBot);
    const bool isHorizontal = horizontal >= vertical;
    const float tsx = 1.0f / (float)im.W, tsy = 1.0f / (float)im.H;  // computeTexelSize
    float pixelStep = isHorizontal ? tsy : tsx;
    const float positiveLuma = isHorizontal ? lumaTop : lumaRight;
    const float negativeLuma = isHorizontal ? lumaBot : lumaLeft;
    const float positiveGradient = fabsf(positiveLuma - lumaCenter);
    const float negativeGradient = fabsf(negativeLuma - lumaCenter);
    float gradient, oppositeLuma;
    if (positiveGradient < negativeGradient) {
        pixelStep = -pixelStep;
        oppositeLuma = negativeLuma;
        gradient = negativeGradient;
    } else {
        oppositeLuma = positiveLuma;
        gradient = positiveGradient;
    }
    float ue = u, ve = v, esx, esy;
    if (isHorizontal) {
        ve += pixelStep * 0.5f;
        esx = tsx;
        esy = 0.0f;
    } else {
        ue += pixelStep * 0.5f;
        esx = 0.0f;
        esy = tsy;
    }
    const float edgeLumaAvg = (lumaCenter + oppositeLuma) * 0.5f;
    const float gradientThreshold = gradient * 0.25f;
    // edgeStepOffsets = 1, 1.5, 2 x 7, 4; a guess jump of 8 when no end is found
    float dist[2], deltas[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const float sgn = i == 0 ? 1.0f : -1.0f;
        const float sx = esx * sgn, sy = esy * sgn;
        float pu = ue + sx * 1.0f, pv = ve + sy * 1.0f;
        float delta = lumaLinear(im, lut, pu, pv) - edgeLumaAvg;
        bool atEnd = fabsf(delta) >= gradientThreshold;
        for (int j = 0; j < 9 && !atEnd; ++j) {
            const float off = j == 0 ? 1.5f : (j == 8 ? 4.0f : 2.0f);
            pu += sx * off;
            pv += sy * off;
            delta = lumaLinear(im, lut, pu, pv) - edgeLumaAvg;
            atEnd = fabsf(delta) >= gradientThreshold;
        }
        if (!atEnd) {
            pu += sx * 8.0f;
            pv += sy * 8.0f;
        }
        deltas[i] = delta;
        dist[i] = isHorizontal ? (sgn * pu) + (-sgn * u) : (sgn * pv) + (-sgn * v);
    }
    float shortest;
    bool deltaSign;
    if (dist[0] <= dist[1]) {
        shortest = dist[0];
        deltaSign = deltas[0] >= 0.0f;
    } else {
        shortest = dist[1];
        deltaSign = deltas[1] >= 0.0f;
    }
    const float edgeBlendFactor =
        deltaSign == (lumaCenter - edgeLumaAvg >= 0.0f) ? 0.0f : 0.5f - shortest / (dist[0] + dist[1]);
    const float finalBlend = fmaxf(pixelBlendFactor, edgeBlendFactor);
    float fu = u, fv = v;
    if (isHorizontal) fv += pixelStep * finalBlend;
    else fu += pixelStep * finalBlend;
    // the final LINEAR tap of the rgb (the luminance target's rgb = the input's codes)
    const float gu = fu * (float)im.W - 0.5f, gv = fv * (float)im.H - 0.5f;
    const float i0f = floorf(gu), j0f = floorf(gv);
    const float a = gu - i0f, b = gv - j0f;
    const int ia = texelIndex(i0f, im.W), ib = texelIndex(i0f + 1.0f, im.W);
    const int ja = texelIndex(j0f, im.H), jb = texelIndex(j0f + 1.0f, im.H);
    const uint32_t t00 = fxaaTexel(im, ia, ja), t10 = fxaaTexel(im, ib, ja);
    const uint32_t t01 = fxaaTexel(im, ia, jb), t11 = fxaaTexel(im, ib, jb);
    uint32_t o = 0xFF000000u;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float val = (1.0f - a) * (1.0f - b) * unorm8(lut, t00, c) + a * (1.0f - b) * unorm8(lut, t10, c) +
                          (1.0f - a) * b * unorm8(lut, t01, c) + a * b * unorm8(lut, t11, c);
        o |= store8(val) << (8 * c);
    }
    out[p] = o;
}

}  // namespace

namespace sgi_launch {

hipError_t taaHistory(const float* taa, long n, float* prev, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const long n3 = 3 * n;
    hipLaunchKernelGGL(taaHistoryKernel, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s, taa, n3, prev);
    return hipGetLastError();
}

hipError_t fxaa(const uint32_t* in, int W, int H, float contrastThreshold, float relativeThreshold,
                float subpixelBlending, uint32_t* out, hipStream_t s) {
    if (W <= 0 || H <= 0) return hipSuccess;
    const dim3 grid((unsigned)((W + 15) / 16), (unsigned)((H + 15) / 16));
    hipLaunchKernelGGL(fxaaKernel, grid, dim3(256), 0, s, FxaaImage{in, W, H}, contrastThreshold, relativeThreshold,
                       subpixelBlending, out);
    return hipGetLastError();
}

hipError_t tonemap(const float* lit, long n, float gamma, uint32_t* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const float invGamma = 1.0f / gamma;  // 1.0 / gamma (:113)
    hipLaunchKernelGGL(tonemapKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, lit, n, invGamma, out);
    return hipGetLastError();
}

}  // namespace sgi_launch
