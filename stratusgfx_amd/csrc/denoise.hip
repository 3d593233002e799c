// A-trous spatiotemporal denoiser — vpl_pbr_gi_denoise.fs:142-435 run as the
// reference's 4-pass loop (StratusRendererBackend.cpp:1912-1951).
//
// Only the reservoir changes between passes (the colour / GI attachment passes
// through until the final pass), so the passes exchange one float4 plane:
//   pass 0  denoiseMergeKernel: centre-only reservoir merge (computeMergedReservoir
//           with minmaxNearest = 0) and normalisation; it also stores each pixel's
//           decoded normal and depth as one float4, the 16 B every later tap reads;
//   pass 1,2 denoiseFilterKernel: 5x5 taps at stride 2^pass with weight
//           pow(max(0, n.n'), 128) * exp(-|z - z'|) (centre weight 1);
//   pass 3  denoiseFilterKernel at stride 8 into the reservoir output, then
//           denoiseTemporalKernel: the temporal blend with the reprojected
//           history and the combine screen + albedo * gi.
//
// Blocks cover 64x4 pixels and are numbered so that each XCD walks a
// contiguous band of rows: the 5x5 taps at stride 8 reach 16 rows up and down,
// and that window stays in the XCD's L2 instead of being fetched once per XCD.
// The edge-stopping weights and the normalisation are continuous, so they use
// hardware exp2 / rcp; the merge test and the temporal similarity test are
// decisions and keep correctly rounded arithmetic.
#include "launch.h"
#include "views.h"

namespace {

using namespace sgi_dev;

__device__ __forceinline__ int clampx(int x, int W) { return x < 0 ? 0 : (x >= W ? W - 1 : x); }

// 64x4-pixel block b of a row-band numbering.
__device__ __forceinline__ void blockXY(int W, int b, int& x, int& y) {
    const int bx = (W + 63) / 64;
    x = (b % bx) * 64 + (int)threadIdx.x % 64;
    y = (b / bx) * 4 + (int)threadIdx.x / 64;
}

// The block of this launch in an XCD-contiguous numbering: each XCD walks a band of rows.
__device__ __forceinline__ void blockPixel(int W, int H, int& x, int& y) {
    const int bx = (W + 63) / 64, by = (H + 3) / 4;
    blockXY(W, (int)xcdSwizzle(blockIdx.x, (uint32_t)(bx * by)), x, y);
}

// Batched launches: the view of this block and the block's pixel in it.
__device__ __forceinline__ const sgi_batch::ViewDesc& viewBlock(const sgi_batch::ViewDesc* views, int numViews,
                                                               int total, int& x, int& y) {
    const int gb = (int)xcdSwizzle(blockIdx.x, (uint32_t)total);
    const sgi_batch::ViewDesc& d = views[sgi_batch::findView<sgi_batch::kDenoiseBlocks>(views, numViews, gb)];
    blockXY(d.g.W, gb - d.dblockStart, x, y);
    return d;
}

// GL LINEAR + CLAMP_TO_EDGE at an arbitrary uv: the four texel indices and
// weights, computed once for every channel read there; blend() is
// ((w00 t00 + w10 t10) + w01 t01) + w11 t11 with w00 = (1 - a)(1 - b),
// w10 = a (1 - b), w01 = (1 - a) b, w11 = a b.
struct BilinTaps {
    size_t q00, q10, q01, q11;
    float w00, w10, w01, w11;
};
__device__ __forceinline__ BilinTaps bilinTaps(int W, int H, float u, float v) {
    const float fu = u * (float)W - 0.5f, fv = v * (float)H - 0.5f;
    const float i0f = floorf(fu), j0f = floorf(fv);
    const float a = fu - i0f, b = fv - j0f;
    const int ia = texelIndex(i0f, W), ib = texelIndex(i0f + 1.0f, W);
    const int ja = texelIndex(j0f, H), jb = texelIndex(j0f + 1.0f, H);
    return BilinTaps{(size_t)ja * W + ia,       (size_t)ja * W + ib,  (size_t)jb * W + ia,
                     (size_t)jb * W + ib,       (1.0f - a) * (1.0f - b), a * (1.0f - b),
                     (1.0f - a) * b,            a * b};
}
__device__ __forceinline__ float blend(const BilinTaps& t, float v00, float v10, float v01, float v11) {
    return t.w00 * v00 + t.w10 * v10 + t.w01 * v01 + t.w11 * v11;
}
__device__ __forceinline__ f3 blend3(const BilinTaps& t, f3 v00, f3 v10, f3 v01, f3 v11) {
    return f3{blend(t, v00.x, v10.x, v01.x, v11.x), blend(t, v00.y, v10.y, v01.y, v11.y),
              blend(t, v00.z, v10.z, v01.z, v11.z)};
}

__device__ __forceinline__ float nearestTap(const float* t, int W, int H, int C, float u, float v, int c) {
    const int i = texelIndex(floorf(u * (float)W), W), j = texelIndex(floorf(v * (float)H), H);
    return t[((size_t)j * W + i) * C + c];
}

// A tile list (int32 [1 + n]: the count, then 16x16 tile indices, row-major
// over ceil(W / 16) tiles per row): the blocks of a fixed grid take its tiles
// in turn, one pixel per thread; an empty list costs one load per block.
constexpr int kListGrid = 1024;
template <class F>
__device__ __forceinline__ void forListedTiles(const int32_t* __restrict__ list, int W, F body) {
    const int n = list[0], tilesX = (W + 15) >> 4;
    for (int i = (int)blockIdx.x; i < n; i += (int)gridDim.x) {
        const int t = list[1 + i];
        body((t % tilesX) * 16 + (int)threadIdx.x % 16, (t / tilesX) * 16 + (int)threadIdx.x / 16);
    }
}

// flags (per 16x16 tile, may be NULL): whether pixel (x, y)'s tile is marked.
__device__ __forceinline__ bool flagged(const uint8_t* flags, int W, int x, int y) {
    return flags && flags[(y >> 4) * ((W + 15) >> 4) + (x >> 4)];
}

__device__ __forceinline__ void mergeBody(const GbufArgs& g, const FrameArgs& fr, const int32_t* __restrict__ visible,
                                          const float4* __restrict__ resIn, float4* __restrict__ resOut,
                                          float4* __restrict__ normals, int x, int y) {
    if (x >= g.W || y >= g.H) return;
    const size_t p = (size_t)y * g.W + x;
    const f3 centerNormal = gNormal(g, p);
    const float centerDepth = g.depth[p];
    normals[p] = make_float4(centerNormal.x, centerNormal.y, centerNormal.z, centerDepth);
    const int numVisible = visible[0];
    float cr0 = 0.0f, cr1 = 0.0f, cr2 = 0.0f, cr3 = 0.0f;
    const float depthCutoff = 0.1f * centerDepth;
    float runningSum = 0.0f;
    const float probabilisticWeight = 1.0f / (float)numVisible;
    bool skip = dot(centerNormal, centerNormal) < 0.906f;  // the centre tap's normal test
    if (!skip) skip = fabsf(centerDepth - centerDepth) > depthCutoff;
    if (!skip) {
        const float4 rr = resIn[p];
        const float randUpdate = random3((float)x + 0.5f, (float)y + 0.5f, fr.time);
        const float probability = rr.w * probabilisticWeight;
        cr3 += probability;
        if (randUpdate < probability / cr3) {
            cr0 += rr.x;
            cr1 += rr.y;
            cr2 += rr.z;
            runningSum += rr.w;
        }
    }
    cr3 = runningSum;
    const float inv = frcp(gmax(kEPS, cr3));
    resOut[p] = make_float4(cr0 * inv, cr1 * inv, cr2 * inv, 1.0f);
}

__global__ __launch_bounds__(256) void denoiseMergeKernel(GbufArgs g, FrameArgs fr, const int32_t* __restrict__ visible,
                                                          const float4* __restrict__ resIn, float4* __restrict__ resOut,
                                                          float4* __restrict__ normals) {
    int x, y;
    blockPixel(g.W, g.H, x, y);
    mergeBody(g, fr, visible, resIn, resOut, normals, x, y);
}

__global__ __launch_bounds__(256) void denoiseMergeListKernel(GbufArgs g, FrameArgs fr,
                                                              const int32_t* __restrict__ visible,
                                                              const float4* __restrict__ resIn,
                                                              float4* __restrict__ resOut, float4* __restrict__ normals,
                                                              const int32_t* __restrict__ list) {
    forListedTiles(list, g.W, [&](int x, int y) { mergeBody(g, fr, visible, resIn, resOut, normals, x, y); });
}

__global__ __launch_bounds__(256) void denoiseMergeViewsKernel(const sgi_batch::ViewDesc* __restrict__ views,
                                                               int numViews, int total) {
    int x, y;
    const sgi_batch::ViewDesc& d = viewBlock(views, numViews, total, x, y);
    mergeBody(d.g, d.fr, d.visible, (const float4*)d.reservoir, d.resA, d.normals, x, y);
}

__device__ __forceinline__ void filterBody(const GbufArgs& g, int stride, const float4* __restrict__ resIn,
                                           const float4* __restrict__ normals, float4* __restrict__ resOut, int x,
                                           int y) {
    if (x >= g.W || y >= g.H) return;
    const int W = g.W, H = g.H;
    const size_t p = (size_t)y * W + x;
    const float4 cn4 = normals[p];
    const f3 centerNormal = f3{cn4.x, cn4.y, cn4.z};
    const float centerDepth = cn4.w;
    float rf0 = 0.0f, rf1 = 0.0f, rf2 = 0.0f, rf3 = 0.0f;
#pragma unroll
    for (int dx = -2; dx <= 2; ++dx) {
        const int tx = clampx(x + dx * stride, W);
#pragma unroll
        for (int dy = -2; dy <= 2; ++dy) {
            const int ty = clampx(y + dy * stride, H);
            const size_t q = (size_t)ty * W + tx;
            const float4 rr = resIn[q];
            float filtered;
            if (dx == 0 && dy == 0) {
                filtered = 1.0f;
            } else {  // filterInput, vpl_pbr_gi_denoise.fs:142-190
                const float4 n4 = normals[q];
                const float wz = fexp(-fabsf(centerDepth - n4.w));
                float wn = gmax(0.0f, dotF(centerNormal, f3{n4.x, n4.y, n4.z}));
                wn = pow128(wn);
                filtered = wn * wz;
            }
            rf0 = __builtin_fmaf(filtered, rr.x, rf0);
            rf1 = __builtin_fmaf(filtered, rr.y, rf1);
            rf2 = __builtin_fmaf(filtered, rr.z, rf2);
            rf3 = __builtin_fmaf(filtered, rr.w, rf3);
        }
    }
    const float inv = frcp(gmax(kEPS, rf3));
    resOut[p] = make_float4(rf0 * inv, rf1 * inv, rf2 * inv, 1.0f);
}

__global__ __launch_bounds__(256) void denoiseFilterKernel(GbufArgs g, int stride, const float4* __restrict__ resIn,
                                                           const float4* __restrict__ normals,
                                                           float4* __restrict__ resOut) {
    int x, y;
    blockPixel(g.W, g.H, x, y);
    filterBody(g, stride, resIn, normals, resOut, x, y);
}

__global__ __launch_bounds__(256) void denoiseFilterListKernel(GbufArgs g, int stride, const float4* __restrict__ resIn,
                                                               const float4* __restrict__ normals,
                                                               float4* __restrict__ resOut,
                                                               const int32_t* __restrict__ list) {
    forListedTiles(list, g.W, [&](int x, int y) { filterBody(g, stride, resIn, normals, resOut, x, y); });
}

// pass 1: resA -> resB, pass 2: resB -> resA, pass 3: resA -> reservoirOut
__global__ __launch_bounds__(256) void denoiseFilterViewsKernel(const sgi_batch::ViewDesc* __restrict__ views,
                                                                int numViews, int total, int pass) {
    int x, y;
    const sgi_batch::ViewDesc& d = viewBlock(views, numViews, total, x, y);
    const float4* in = pass == 2 ? d.resB : d.resA;
    float4* out = pass == 1 ? d.resB : (pass == 2 ? d.resA : (float4*)d.reservoirOut);
    filterBody(d.g, 1 << pass, in, d.normals, out, x, y);
}

// The final pass's temporal blend and combine (vpl_pbr_gi_denoise.fs:336-434) on
// the filtered reservoir the last filter pass wrote.  FROM_NET: the filtered
// value is the learned denoiser's output (net, FP32 [h][w][3]), written back
// as the output reservoir (net, 1), and the centre's decoded normal and depth
// come from the G-buffer (the same values the merge pass stores).
template <bool FROM_NET = false>
__device__ __forceinline__ void temporalBody(const GbufArgs& g, const GbufArgs& prev, const FrameArgs& fr,
                                             const float4* __restrict__ filtered, const float4* __restrict__ normals,
                                             const float* __restrict__ screen, const float* __restrict__ color,
                                             const float* __restrict__ prevGi, const float* __restrict__ prevHistory,
                                             float* __restrict__ combined, float* __restrict__ giOut,
                                             float* __restrict__ historyOut, int x, int y,
                                             const float* __restrict__ net = nullptr,
                                             float4* __restrict__ reservoirOut = nullptr) {
    if (x >= g.W || y >= g.H) return;
    const int W = g.W, H = g.H;
    const size_t p = (size_t)y * W + x;
    f3 centerNormal, shadowFactor;
    float centerDepth;
    if constexpr (FROM_NET) {
        centerNormal = gNormal(g, p);
        centerDepth = g.depth[p];
        shadowFactor = ld3(net + 3 * p);
        reservoirOut[p] = make_float4(shadowFactor.x, shadowFactor.y, shadowFactor.z, 1.0f);
    } else {
        const float4 cn4 = normals[p];
        centerNormal = f3{cn4.x, cn4.y, cn4.z};
        centerDepth = cn4.w;
        const float4 sf4 = filtered[p];
        shadowFactor = f3{sf4.x, sf4.y, sf4.z};
    }
    const f3 gi = ld3(color + 3 * p);
    float historyAccum = prevHistory ? prevHistory[p] : 0.0f;
    float accumMultiplier = 1.0f;
    const f3 currGi = shadowFactor;
    const float u = texelCentre(x, W), v = texelCentre(y, H);
    const float2 vel = gVelocity(g, p);
    const float pu = u - vel.x, pv = v - vel.y;
    // the previous frame's depth, normal and GI at uv - velocity: one set of taps, each texel read once
    const BilinTaps bt = bilinTaps(W, H, pu, pv);
    const float prevCenterDepth =
        blend(bt, prev.depth[bt.q00], prev.depth[bt.q10], prev.depth[bt.q01], prev.depth[bt.q11]);
    const f3 prevCenterNormal = normalize(blend3(bt, gNormalRaw(prev, bt.q00), gNormalRaw(prev, bt.q10),
                                                 gNormalRaw(prev, bt.q01), gNormalRaw(prev, bt.q11)) *
                                              2.0f -
                                          splat(1.0f));
    f3 prevGiV = splat(0.0f);
    if (prevGi)
        prevGiV = blend3(bt, ld3(prevGi + 3 * bt.q00), ld3(prevGi + 3 * bt.q10), ld3(prevGi + 3 * bt.q01),
                         ld3(prevGi + 3 * bt.q11));
    const float currId = g.id ? g.id[p] : 0.0f;
    const float prevId = prev.id ? nearestTap(prev.id, W, H, 1, pu, pv, 0) : 0.0f;
    const float wn = gmax(0.0f, dot(centerNormal, prevCenterNormal));
    const float wz = expf(-fabsf(centerDepth - prevCenterDepth));
    const float wid = currId != prevId ? 0.0f : 1.0f;
    const float similarity = wn * wz * wid;
    if (similarity < 0.99f) accumMultiplier = 0.0f;
    historyAccum = gmin(1.0f + historyAccum * accumMultiplier, fr.fps);
    const float maxAccumulationFactor = frcp(historyAccum);
    const f3 illumAvg = mixv(prevGiV, currGi, maxAccumulationFactor);
    const f3 sc = ld3(screen + 3 * p);
    const f3 comb = sc + gi * illumAvg;
    combined[3 * p + 0] = comb.x;
    combined[3 * p + 1] = comb.y;
    combined[3 * p + 2] = comb.z;
    giOut[3 * p + 0] = illumAvg.x;
    giOut[3 * p + 1] = illumAvg.y;
    giOut[3 * p + 2] = illumAvg.z;
    historyOut[p] = historyAccum;
}

__global__ __launch_bounds__(256) void denoiseTemporalKernel(GbufArgs g, GbufArgs prev, FrameArgs fr,
                                                             const float4* __restrict__ filtered,
                                                             const float4* __restrict__ normals,
                                                             const float* __restrict__ screen,
                                                             const float* __restrict__ color,
                                                             const float* __restrict__ prevGi,
                                                             const float* __restrict__ prevHistory,
                                                             float* __restrict__ combined, float* __restrict__ giOut,
                                                             float* __restrict__ historyOut) {
    int x, y;
    blockPixel(g.W, g.H, x, y);
    temporalBody(g, prev, fr, filtered, normals, screen, color, prevGi, prevHistory, combined, giOut, historyOut, x, y);
}

// The learned denoiser's temporal pass when no tile falls back to the a-trous filter.
__global__ __launch_bounds__(256) void denoiseTemporalNetKernel(GbufArgs g, GbufArgs prev, FrameArgs fr,
                                                                const float* __restrict__ net,
                                                                const float* __restrict__ screen,
                                                                const float* __restrict__ color,
                                                                const float* __restrict__ prevGi,
                                                                const float* __restrict__ prevHistory,
                                                                float* __restrict__ combined,
                                                                float* __restrict__ giOut,
                                                                float4* __restrict__ reservoirOut,
                                                                float* __restrict__ historyOut,
                                                                const uint8_t* __restrict__ fallback,
                                                                const float4* __restrict__ normals) {
    int x, y;
    blockPixel(g.W, g.H, x, y);
    if (flagged(fallback, g.W, x, y)) {  // the a-trous result already in reservoirOut, its decoded normals
        temporalBody(g, prev, fr, reservoirOut, normals, screen, color, prevGi, prevHistory, combined, giOut,
                     historyOut, x, y);
        return;
    }
    temporalBody<true>(g, prev, fr, nullptr, nullptr, screen, color, prevGi, prevHistory, combined, giOut, historyOut,
                       x, y, net, reservoirOut);
}

// Batched views: the previous frame's G-buffer is the current one (a single frame per view).
__global__ __launch_bounds__(256) void denoiseTemporalViewsKernel(const sgi_batch::ViewDesc* __restrict__ views,
                                                                  int numViews, int total) {
    int x, y;
    const sgi_batch::ViewDesc& d = viewBlock(views, numViews, total, x, y);
    temporalBody(d.g, d.g, d.fr, (const float4*)d.reservoirOut, d.normals, d.direct, d.color, d.prevGi,
                 d.prevHistory, d.combined, d.giOut, d.historyOut, x, y);
}

// taa.fs:12-100 (PerformTaaPostFx_, StratusRendererBackend.cpp:2326-2352): the
// velocity texel, the current colour and its 4 neighbours at texel centres,
// the previous frame bilinearly at uv - velocity, the history clamped to the
// neighbourhood's range in tonemapped space (aa_common.glsl: c / (c + 1)),
// both mapped back (c / (1 - c)) and mixed 0.1 / 0.9.  Per pixel: 5 current
// taps, 4 history taps, one velocity texel; FP32 in, FP32 out.
__device__ __forceinline__ float aaTonemap(float c) { return c / (c + 1.0f); }
__device__ __forceinline__ float aaInverse(float c) { return c / (1.0f - c); }

__device__ __forceinline__ f3 texel3(const float* t, int W, int x, int y) {
    const float* q = t + ((size_t)y * W + x) * 3;
    return f3{q[0], q[1], q[2]};
}

__global__ __launch_bounds__(256) void taaKernel(const float* __restrict__ screen, const float* __restrict__ prev,
                                                  const float* __restrict__ velocity, int W, int H,
                                                  float* __restrict__ out) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long)W * H) return;
    const int x = (int)(p % W), y = (int)(p / W);
    const float u = texelCentre(x, W), v = texelCentre(y, H);
    const float vu = velocity ? velocity[2 * p] : 0.0f, vv = velocity ? velocity[2 * p + 1] : 0.0f;
    // bilinear history coordinates (GL LINEAR, CLAMP_TO_EDGE), shared by the three channels
    const float fu = (u - vu) * (float)W - 0.5f, fv = (v - vv) * (float)H - 0.5f;
    const float i0f = floorf(fu), j0f = floorf(fv);
    const float a = fu - i0f, b = fv - j0f;
    const int ia = texelIndex(i0f, W), ib = texelIndex(i0f + 1.0f, W);
    const int ja = texelIndex(j0f, H), jb = texelIndex(j0f + 1.0f, H);
    const f3 h00 = texel3(prev, W, ia, ja), h10 = texel3(prev, W, ib, ja);
    const f3 h01 = texel3(prev, W, ia, jb), h11 = texel3(prev, W, ib, jb);
    const f3 c0 = texel3(screen, W, x, y);
    const f3 c1 = texel3(screen, W, x, min(y + 1, H - 1)), c2 = texel3(screen, W, x, max(y - 1, 0));
    const f3 c3 = texel3(screen, W, min(x + 1, W - 1), y), c4 = texel3(screen, W, max(x - 1, 0), y);
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    const float mixA = 0.9f;
    const float h00c[3] = {h00.x, h00.y, h00.z}, h10c[3] = {h10.x, h10.y, h10.z};
    const float h01c[3] = {h01.x, h01.y, h01.z}, h11c[3] = {h11.x, h11.y, h11.z};
    const float cc[5][3] = {{c0.x, c0.y, c0.z}, {c1.x, c1.y, c1.z}, {c2.x, c2.y, c2.z}, {c3.x, c3.y, c3.z},
                            {c4.x, c4.y, c4.z}};
    for (int c = 0; c < 3; ++c) {
        const float cur = cc[0][c];
        const float hist = w00 * h00c[c] + w10 * h10c[c] + w01 * h01c[c] + w11 * h11c[c];
        const float prevColor = aaTonemap(hist);
        const float mn = aaTonemap(fminf(cur, fminf(cc[1][c], fminf(cc[2][c], fminf(cc[3][c], cc[4][c])))));
        const float mx = aaTonemap(fmaxf(cur, fmaxf(cc[1][c], fmaxf(cc[2][c], fmaxf(cc[3][c], cc[4][c])))));
        const float curT = aaTonemap(cur);
        const float prevC = fminf(fmaxf(prevColor, mn), mx);
        out[3 * p + c] = aaInverse(curT) * (1.0f - mixA) + aaInverse(prevC) * mixA;
    }
}

}  // namespace

namespace sgi_launch {

hipError_t taa(const float* screen, const float* prev, const float* velocity, int W, int H, float* out, hipStream_t s) {
    const long n = (long)W * H;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(taaKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, screen, prev, velocity, W, H, out);
    return hipGetLastError();
}

hipError_t denoiseSpatial(const GbufArgs& cur, const FrameArgs& fr, const int32_t* visible, const float* reservoir,
                          float* reservoirOut, const DenoiseBuffers& ws, const int32_t* tiles, bool mergeOnly,
                          hipStream_t s) {
    const dim3 grid(((cur.W + 63) / 64) * ((cur.H + 3) / 4)), block(256), lgrid(kListGrid);
    if (tiles)
        hipLaunchKernelGGL(denoiseMergeListKernel, lgrid, block, 0, s, cur, fr, visible, (const float4*)reservoir,
                           ws.resA, ws.normals, tiles);
    else
        hipLaunchKernelGGL(denoiseMergeKernel, grid, block, 0, s, cur, fr, visible, (const float4*)reservoir, ws.resA,
                           ws.normals);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || mergeOnly) return e;
    float4* const in[3] = {ws.resA, ws.resB, ws.resA};
    float4* const out[3] = {ws.resB, ws.resA, (float4*)reservoirOut};
    for (int pass = 0; pass < 3 && e == hipSuccess; ++pass) {
        if (tiles)
            hipLaunchKernelGGL(denoiseFilterListKernel, lgrid, block, 0, s, cur, 2 << pass, in[pass], ws.normals,
                               out[pass], tiles);
        else
            hipLaunchKernelGGL(denoiseFilterKernel, grid, block, 0, s, cur, 2 << pass, in[pass], ws.normals,
                               out[pass]);
        e = hipGetLastError();
    }
    return e;
}

hipError_t denoiseTemporal(const GbufArgs& cur, const GbufArgs& prev, const FrameArgs& fr, const float* screen,
                           const float* color, const float* filtered, const float* prevGi, const float* prevHistory,
                           float* combined, float* giOut, float* historyOut, const DenoiseBuffers& ws, hipStream_t s) {
    hipLaunchKernelGGL(denoiseTemporalKernel, dim3(((cur.W + 63) / 64) * ((cur.H + 3) / 4)), dim3(256), 0, s, cur,
                       prev, fr, (const float4*)filtered, ws.normals, screen, color, prevGi, prevHistory, combined,
                       giOut, historyOut);
    return hipGetLastError();
}

hipError_t denoiseTemporalNet(const GbufArgs& cur, const GbufArgs& prev, const FrameArgs& fr, const float* screen,
                              const float* color, const float* net, const float* prevGi, const float* prevHistory,
                              float* combined, float* giOut, float* reservoirOut, float* historyOut,
                              const uint8_t* fallback, const float4* normals, hipStream_t s) {
    const dim3 grid(((cur.W + 63) / 64) * ((cur.H + 3) / 4)), block(256);
    hipLaunchKernelGGL(denoiseTemporalNetKernel, grid, block, 0, s, cur, prev, fr, net, screen, color, prevGi,
                       prevHistory, combined, giOut, (float4*)reservoirOut, historyOut, fallback, normals);
    return hipGetLastError();
}

hipError_t denoise(const GbufArgs& cur, const GbufArgs& prev, const FrameArgs& fr, const int32_t* visible,
                   const float* screen, const float* color, const float* reservoir, const float* prevGi,
                   const float* prevHistory, float* combined, float* giOut, float* reservoirOut, float* historyOut,
                   const DenoiseBuffers& ws, hipStream_t s) {
    const hipError_t e = denoiseSpatial(cur, fr, visible, reservoir, reservoirOut, ws, nullptr, false, s);
    if (e != hipSuccess) return e;
    return denoiseTemporal(cur, prev, fr, screen, color, reservoirOut, prevGi, prevHistory, combined, giOut,
                           historyOut, ws, s);
}

hipError_t denoiseViews(const sgi_batch::ViewDesc* views, int numViews, int totalBlocks, hipStream_t s) {
    hipLaunchKernelGGL(denoiseMergeViewsKernel, dim3(totalBlocks), dim3(256), 0, s, views, numViews, totalBlocks);
    hipError_t e = hipGetLastError();
    for (int pass = 1; pass <= 3 && e == hipSuccess; ++pass) {
        hipLaunchKernelGGL(denoiseFilterViewsKernel, dim3(totalBlocks), dim3(256), 0, s, views, numViews, totalBlocks,
                           pass);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(denoiseTemporalViewsKernel, dim3(totalBlocks), dim3(256), 0, s, views, numViews, totalBlocks);
    return hipGetLastError();
}

}  // namespace sgi_launch
