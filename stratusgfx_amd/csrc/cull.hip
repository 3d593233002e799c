// VPL culling kernels.
//
// cullVisibilityKernel + cullCompactKernel — viscull_vpls.cs:47-120: per VPL,
// the sun-visibility test through the cascades, then a stable compaction (wave
// ballot + block prefix) of the visible records in ascending input order.
//
// cullTilesKernel — vpl_tiled_deferred_culling_stage1.cs:50-90 and
// _stage2.cs:61-129 fused: one wave per tile.  Stage 1 reconstructs the
// tile's pixels in parallel into LDS and then sums each of the six components
// serially in row-major local order, which is the reference's summation order
// (thread 0's loop), so the tile mean is bit-identical.  Stage 2 evaluates 64
// VPLs per step across the lanes and keeps the sorted top-K list in lane
// registers (lane k holds entry k); candidates below the current K-th ratio
// are inserted one at a time in slot order with a ballot for the position and
// a lane shift, which reproduces the reference's strict-< insertion sort.
//
// colorVplsKernel — vpl_light_color.cs:41-76: one thread per visible slot
// reads the diffuse and distance texels in direction -sun and writes the
// slot's colour and specular position.
#include <hip/hip_fp16.h>

#include "launch.h"
#include "views.h"

namespace {

using namespace sgi_dev;

constexpr int kCullThreads = 1024;

// Per-VPL sun visibility, one thread per VPL over many blocks (each VPL's two
// 4x4 cascade texel blocks are scattered; spreading them over CUs keeps the
// loads in flight on many L1s).
__device__ __forceinline__ void cullVisibilityBody(const CsmArgs& csm, f3 sun, const sgi_vpl* __restrict__ in, int n,
                                                   uint8_t* __restrict__ flags, int i) {
    if (i >= n) return;
    const f3 p = ld3(in[i].position);
    const f3 blends = cascadeBlends(csm, p);
    // the 4x4-block form of the 36 cascade compares (one read per texel; the same value)
    const float shadowFactor = 1.0f - infiniteShadowValueBlock(csm, sun, p, blends, sun, 0.0f);
    flags[i] = shadowFactor < 1.0f ? 1 : 0;
}

__global__ __launch_bounds__(64) void cullVisibilityKernel(CsmArgs csm, f3 sun, const sgi_vpl* __restrict__ in, int n,
                                                          uint8_t* __restrict__ flags) {
    cullVisibilityBody(csm, sun, in, n, flags, blockIdx.x * 64 + threadIdx.x);
}

// Batched: blocks [v * bpv, (v + 1) * bpv) test the VPLs against view v's cascades.
__global__ __launch_bounds__(64) void cullVisibilityViewsKernel(const sgi_batch::ViewDesc* __restrict__ views,
                                                               const sgi_vpl* __restrict__ in, int n, int bpv) {
    const sgi_batch::ViewDesc& d = views[blockIdx.x / bpv];
    cullVisibilityBody(d.csm, ld3(d.fr.sunDir), in, n, d.visFlags, (blockIdx.x % bpv) * 64 + threadIdx.x);
}

// Stable compaction of the flags: wave ballots and a block prefix, in input
// order. The block takes kCompactRounds chunks of kCullThreads items at a
// time: the chunks' flags are loaded together before the one barrier, then
// positions follow chunk, wave and lane order (each chunk's items after all
// earlier chunks') and the visible records are copied. Few registers, so the
// block is placed beside a large kernel running on another stream.
constexpr int kCompactRounds = 4;
__device__ __forceinline__ void cullCompactBody(const sgi_vpl* __restrict__ in, int n, const uint8_t* __restrict__ flags,
                                                sgi_vpl* __restrict__ out, int32_t* __restrict__ visible,
                                                int (*waveCount)[kCullThreads / 64]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    int running = 0;
    for (int c0 = 0; c0 < n; c0 += kCullThreads * kCompactRounds) {
        uint8_t f[kCompactRounds];
#pragma unroll
        for (int k = 0; k < kCompactRounds; ++k) {
            const int i = c0 + k * kCullThreads + (int)threadIdx.x;
            f[k] = i < n ? flags[i] : 0;
        }
        int lanePrefix[kCompactRounds];
#pragma unroll
        for (int k = 0; k < kCompactRounds; ++k) {
            const uint64_t m = __ballot(f[k] != 0);
            lanePrefix[k] = __popcll(m & below);
            if (lane == 0) waveCount[k][wave] = __popcll(m);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kCompactRounds; ++k) {
            int wavePrefix = 0, total = 0;
            for (int w = 0; w < kCullThreads / 64; ++w) {
                const int c = waveCount[k][w];
                if (w < wave) wavePrefix += c;
                total += c;
            }
            const int pos = running + wavePrefix + lanePrefix[k];
            if (f[k] != 0 && pos < SGI_MAX_TOTAL_VPLS_PER_FRAME) {
                const int i = c0 + k * kCullThreads + (int)threadIdx.x;
                out[pos] = in[i];
                visible[1 + pos] = i;
            }
            running += total;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) visible[0] = running < SGI_MAX_TOTAL_VPLS_PER_FRAME ? running : SGI_MAX_TOTAL_VPLS_PER_FRAME;
}

__global__ __launch_bounds__(kCullThreads) void cullCompactKernel(const sgi_vpl* __restrict__ in, int n,
                                                                  const uint8_t* __restrict__ flags,
                                                                  sgi_vpl* __restrict__ out,
                                                                  int32_t* __restrict__ visible) {
    __shared__ int waveCount[kCompactRounds][kCullThreads / 64];
    cullCompactBody(in, n, flags, out, visible, waveCount);
}

// Batched: one block per view.
__global__ __launch_bounds__(kCullThreads) void cullCompactViewsKernel(const sgi_batch::ViewDesc* __restrict__ views,
                                                                       const sgi_vpl* __restrict__ in, int n) {
    __shared__ int waveCount[kCompactRounds][kCullThreads / 64];
    const sgi_batch::ViewDesc& d = views[blockIdx.x];
    cullCompactBody(in, n, d.visFlags, d.vplUpdated, d.visible, waveCount);
}

__global__ void gatherEntriesKernel(const sgi_atlas_entry* __restrict__ all, const int32_t* __restrict__ visible,
                                    int maxCount, sgi_atlas_entry* __restrict__ out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int count = visible[0];
    if (s < count && s < maxCount) out[s] = all[visible[1 + s]];
}

// vpl_light_color.cs:49-74.  The diffuse and distance atlases have one
// geometry (one framebuffer's two attachments), so both lookups of direction
// -sun share one texel offset.  Built with -ffp-contract=off: every product and
// difference is rounded on its own, in the shader's order.
template <int ZFMT>
__global__ __launch_bounds__(256) void colorVplsKernel(CubeArgs diffuse, CubeArgs dist, f3 sun, f3 sunColor,
                                                        float mult, const int32_t* __restrict__ visible,
                                                        int maxCount, const sgi_atlas_entry* __restrict__ entries,
                                                        sgi_vpl* __restrict__ vpls) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int count = visible[0];
    if (s >= count || s >= maxCount) return;
    const sgi_atlas_entry e = entries[s];
    const f3 d{-sun.x, -sun.y, -sun.z};
    const uint32_t off = cubeTexelOffset(d, diffuse.res);
    const uint32_t rgb = ((const uint32_t*)cubeLayerBase(diffuse, e.index, e.layer, 4))[off];
    const void* zb = cubeLayerBase(dist, e.index, e.layer, texelBytes(ZFMT));
    float z;
    if constexpr (ZFMT == SGI_TEXEL_F32) z = ((const float*)zb)[off];
    else z = decode16<ZFMT>(((const unsigned short*)zb)[off]);
    sgi_vpl& v = vpls[s];
    const float intensity = v.intensity;
    const float scaled = 1.1f * (v.radius * z);  // 1.1 * magnitude (:53, :71)
    const float c[3] = {(float)(rgb & 0xFFu) / 255.0f, (float)((rgb >> 8) & 0xFFu) / 255.0f,
                        (float)((rgb >> 16) & 0xFFu) / 255.0f};
    const float sc[3] = {sunColor.x, sunColor.y, sunColor.z}, sd[3] = {sun.x, sun.y, sun.z};
    for (int k = 0; k < 3; ++k) {
        v.color[k] = c[k] * sc[k] * intensity * mult;  // color / totalColors (1) * intensity * colorMultiplier (:73)
        v.specularPosition[k] = v.position[k] - scaled * sd[k];
    }
    v.color[3] = 1.0f;
    v.specularPosition[3] = 1.0f;
}

// Batched: blocks [v * bpv, (v + 1) * bpv) gather view v's visible slots.
__global__ void gatherEntriesViewsKernel(const sgi_batch::ViewDesc* __restrict__ views,
                                         const sgi_atlas_entry* __restrict__ all, int maxCount, int bpv) {
    const sgi_batch::ViewDesc& d = views[blockIdx.x / bpv];
    const int s = (blockIdx.x % bpv) * blockDim.x + threadIdx.x;
    const int count = d.visible[0];
    if (s < count && s < maxCount) d.entries[s] = all[d.visible[1 + s]];
}

// Floats of LDS a tile-culling block uses: three per-pixel components at a time, later the
// stage-2 histogram (256 counters) and compacted slot list (at least 256 entries). Three rather
// than all six components at once keep 16x16 blocks at 3 KB, which lets a CU hold the eight
// one-wave blocks per SIMD that cover every tile of a 1080p frame in one round.
template <int TS>
constexpr int cullLdsFloats() {
    return 3 * TS * TS > 256 + 128 ? 3 * TS * TS : 256 + 128;
}

template <int TS>
__device__ __forceinline__ void cullTileBody(const GbufArgs& g, const FrameArgs& fr, int t, int tilesX, int K,
                                             const sgi_vpl* __restrict__ vpls, const int32_t* __restrict__ visible,
                                             int remap, sgi_tile_stage1* __restrict__ stage1,
                                             int32_t* __restrict__ tiles, float* comp, float* sums) {
    constexpr int NPX = TS * TS;
    const int tx = t % tilesX, ty = t / tilesX;
    const int lane = threadIdx.x;
    const int x0 = tx * TS, y0 = ty * TS;
    const int vw = min(TS, g.W - x0), vh = min(TS, g.H - y0);

    // Stage 1: per-pixel world positions, then decoded normals, each summed by the reference's
    // serial loop (one component per lane).
    for (int half = 0; half < 2; ++half) {
        for (int l = lane; l < NPX; l += 64) {
            const int lx = l % TS, ly = l / TS;
            if (lx < vw && ly < vh) {
                const int x = x0 + lx, y = y0 + ly;
                const size_t p = (size_t)y * g.W + x;
                const f3 v = half == 0
                                 ? worldPositionFromDepth(texelCentre(x, g.W), texelCentre(y, g.H), g.depth[p], fr.invPV)
                                 : gNormal(g, p);
                comp[l] = v.x; comp[NPX + l] = v.y; comp[2 * NPX + l] = v.z;
            }
        }
        __syncthreads();
        if (lane < 3) {
            float s = 0.0f;
            for (int ly = 0; ly < vh; ++ly)
                for (int lx = 0; lx < vw; ++lx) s += comp[lane * NPX + ly * TS + lx];
            sums[3 * half + lane] = s;
        }
        __syncthreads();
    }
    const float cnt = (float)(vw * vh);
    const f3 P = f3{sums[0], sums[1], sums[2]} / cnt;
    const f3 Nrm = normalize(f3{sums[3], sums[4], sums[5]} / cnt);
    if (stage1 && lane == 0) {
        sgi_tile_stage1 o;
        o.averageLocalPosition[0] = P.x; o.averageLocalPosition[1] = P.y; o.averageLocalPosition[2] = P.z;
        o.averageLocalPosition[3] = 1.0f;
        o.averageLocalNormal[0] = Nrm.x; o.averageLocalNormal[1] = Nrm.y; o.averageLocalNormal[2] = Nrm.z;
        o.averageLocalNormal[3] = 1.0f;
        stage1[t] = o;
    }

    const int numVisible = visible[0];
    // Stage 2a: an upper bound of the K-th smallest candidate ratio, so that stage 2 inserts
    // little beyond the final list (the insertions, not the ratios, dominate its cost). A
    // 256-bucket histogram over [0, 1) of estimated ratios (hardware sqrt and reciprocal: within
    // 2^-21 relative of the correctly rounded ratio) counts only VPLs that are certainly
    // candidates (the exact side test, estimate <= 1 - 2^-18). If K of them fall in buckets up to
    // b, K candidates have ratios below T = (b + 1) / 256 times (1 + 2^-21), so the K-th smallest
    // ratio lies below cap = T (1 + 2^-18): stage 2, started with cap in every slot instead of
    // FLT_MAX, inserts in slot order only the candidates below cap, a subsequence that holds the
    // whole stable top-K, and so ends with the same list. The histogram reuses comp (free once
    // the sums are formed).
    uint32_t* hist = reinterpret_cast<uint32_t*>(comp);
    for (int i = lane; i < 256; i += 64) hist[i] = 0u;
    __syncthreads();
    for (int b = 0; b < numVisible; b += 64) {
        const int s = b + lane;
        if (s < numVisible) {
            const f3 lmf = ld3(vpls[s].position) - P;
            if (!(dot(Nrm, lmf) < 0.0f)) {
                const float r = __builtin_amdgcn_sqrtf(dot(lmf, lmf)) * __builtin_amdgcn_rcpf(vpls[s].radius);
                if (r >= 0.0f && r <= 1.0f - 0x1p-18f) atomicAdd(&hist[(int)(r * 256.0f)], 1u);
            }
        }
    }
    __syncthreads();
    float cap = kFLOAT_MAX;
    {
        const uint32_t h0 = hist[4 * lane], h1 = hist[4 * lane + 1], h2 = hist[4 * lane + 2], h3 = hist[4 * lane + 3];
        const uint32_t mine = h0 + h1 + h2 + h3;
        uint32_t incl = mine;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t v = __shfl_up(incl, o, 64);
            if (lane >= o) incl += v;
        }
        const uint64_t reach = __ballot(incl >= (uint32_t)K);
        if (reach) {
            const int L = __builtin_ctzll(reach);
            const uint32_t e0 = incl - mine;
            const int within = e0 + h0 >= (uint32_t)K ? 0 : (e0 + h0 + h1 >= (uint32_t)K ? 1 : (e0 + h0 + h1 + h2 >= (uint32_t)K ? 2 : 3));
            const int bucket = 4 * L + __builtin_amdgcn_readlane(within, L);
            cap = (float)(bucket + 1) * (1.0f / 256.0f) * (1.0f + 0x1p-18f);
        }
    }

    // Stage 2b: the slots that may lie below cap, compacted in slot order into comp after the
    // histogram (the exact side test, and an estimate within 2^-21 of the ratio kept unless it
    // exceeds cap by more than 2^-19 relative: every candidate below cap is kept). When there is
    // no cap, or more slots than fit, stage 2 scans every slot.
    uint16_t* list = reinterpret_cast<uint16_t*>(hist + 256);
    constexpr int kListCap = (cullLdsFloats<TS>() - 256) * 2;
    int n = numVisible;
    bool useList = false;
    if (cap < kFLOAT_MAX) {
        const float keepBelow = cap * (1.0f + 0x1p-19f);
        int nList = 0;
        for (int b = 0; b < numVisible; b += 64) {
            const int s = b + lane;
            bool keep = false;
            if (s < numVisible) {
                const f3 lmf = ld3(vpls[s].position) - P;
                if (!(dot(Nrm, lmf) < 0.0f))
                    keep = !(__builtin_amdgcn_sqrtf(dot(lmf, lmf)) * __builtin_amdgcn_rcpf(vpls[s].radius) > keepBelow);
            }
            const uint64_t m = __ballot(keep);
            const int pos = nList + __popcll(m & ((1ull << lane) - 1ull));
            if (keep && pos < kListCap) list[pos] = (uint16_t)s;
            nList += __popcll(m);
        }
        useList = nList <= kListCap;
        if (useList) n = nList;
        __syncthreads();
    }

    // Stage 2: sorted top-K in lane registers, over the compacted slots or every slot.
    float myDist = cap;
    int myIdx = 0;
    int count = 0;
    for (int b = 0; b < n; b += 64) {
        const int i = b + lane;
        const int s = i < n ? (useList ? (int)list[i] : i) : 0;
        bool cand = false;
        float ratio = 0.0f;
        if (i < n) {
            const f3 lmf = ld3(vpls[s].position) - P;
            const float side = dot(Nrm, lmf);
            if (!(side < 0.0f)) {
                const float distance = length(lmf);
                ratio = distance / vpls[s].radius;
                cand = !(ratio > 1.0f) && ratio == ratio;
            }
        }
        // K-1 and c are wave-uniform, so the broadcasts are v_readlane
        const float thr = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myDist), K - 1));
        uint64_t m = __ballot(cand && ratio < thr);
        while (m) {
            const int c = __builtin_ctzll(m);
            m &= m - 1;
            const float r = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ratio), c));
            const float thrNow = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myDist), K - 1));
            if (!(r < thrNow)) continue;
            const int p = __popcll(__ballot(lane < K && myDist <= r));
            const int sc = __builtin_amdgcn_readlane(s, c);
            // lane i takes lane i-1's entry: a DPP wave shift, no LDS round trip
            const float upD = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(myDist), __float_as_int(myDist),
                                                                         0x138, 0xF, 0xF, false));
            const int upI = __builtin_amdgcn_update_dpp(myIdx, myIdx, 0x138, 0xF, 0xF, false);
            if (lane < K) {
                if (lane > p) {
                    myDist = upD;
                    myIdx = upI;
                } else if (lane == p) {
                    myDist = r;
                    myIdx = remap ? visible[1 + sc] : sc;
                }
            }
            count = count + 1 < K ? count + 1 : K;
        }
    }
    int32_t* rec = tiles + (size_t)t * (1 + K);
    if (lane == 0) rec[0] = count;
    if (lane < count) rec[1 + lane] = myIdx;
}

template <int TS>
__global__ __launch_bounds__(64) void cullTilesKernel(GbufArgs g, FrameArgs fr, int tilesX, int nTiles, int K,
                                                      const sgi_vpl* __restrict__ vpls,
                                                      const int32_t* __restrict__ visible, int remap,
                                                      sgi_tile_stage1* __restrict__ stage1,
                                                      int32_t* __restrict__ tiles) {
    __shared__ float comp[cullLdsFloats<TS>()];
    __shared__ float sums[6];
    cullTileBody<TS>(g, fr, (int)xcdSwizzle(blockIdx.x, (uint32_t)nTiles), tilesX, K, vpls, visible, remap, stage1,
                     tiles, comp, sums);
}

// Batched: the tiles of every view, numbered view after view.
template <int TS>
__global__ __launch_bounds__(64) void cullTilesViewsKernel(const sgi_batch::ViewDesc* __restrict__ views,
                                                           int numViews, int totalTiles, int K) {
    __shared__ float comp[cullLdsFloats<TS>()];
    __shared__ float sums[6];
    const int gt = (int)xcdSwizzle(blockIdx.x, (uint32_t)totalTiles);
    const sgi_batch::ViewDesc& d = views[sgi_batch::findView<sgi_batch::kTiles>(views, numViews, gt)];
    cullTileBody<TS>(d.g, d.fr, gt - d.tileStart, d.tilesX, K, d.vplUpdated, d.visible, 0, nullptr, d.tiles, comp,
                     sums);
}

}  // namespace

namespace sgi_launch {

hipError_t cullGlobal(const CsmArgs& csm, const FrameArgs& fr, const sgi_vpl* in, int n, sgi_vpl* out,
                      int32_t* visible, uint8_t* flags, hipStream_t s) {
    if (n > 0) {
        hipLaunchKernelGGL(cullVisibilityKernel, dim3((n + 63) / 64), dim3(64), 0, s, csm, ld3(fr.sunDir), in, n, flags);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(cullCompactKernel, dim3(1), dim3(kCullThreads), 0, s, in, n, flags, out, visible);
    return hipGetLastError();
}

hipError_t gatherVisibleEntries(const sgi_atlas_entry* all, const int32_t* visible, int maxCount, sgi_atlas_entry* out,
                                hipStream_t s) {
    if (maxCount <= 0) return hipSuccess;
    hipLaunchKernelGGL(gatherEntriesKernel, dim3((maxCount + 255) / 256), dim3(256), 0, s, all, visible, maxCount, out);
    return hipGetLastError();
}

hipError_t cullTiles(const GbufArgs& g, const FrameArgs& fr, int tileSize, int tilesX, int tilesY, int K,
                     const sgi_vpl* vpls, const int32_t* visible, int remap, sgi_tile_stage1* stage1, int32_t* tiles,
                     hipStream_t s) {
    const int nTiles = tilesX * tilesY;
    switch (tileSize) {
        case 8:
            hipLaunchKernelGGL(cullTilesKernel<8>, dim3(nTiles), dim3(64), 0, s, g, fr, tilesX, nTiles, K, vpls,
                               visible, remap, stage1, tiles);
            break;
        case 16:
            hipLaunchKernelGGL(cullTilesKernel<16>, dim3(nTiles), dim3(64), 0, s, g, fr, tilesX, nTiles, K, vpls,
                               visible, remap, stage1, tiles);
            break;
        case 32:
            hipLaunchKernelGGL(cullTilesKernel<32>, dim3(nTiles), dim3(64), 0, s, g, fr, tilesX, nTiles, K, vpls,
                               visible, remap, stage1, tiles);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t cullGlobalViews(const sgi_batch::ViewDesc* views, int numViews, const sgi_vpl* in, int n, hipStream_t s) {
    if (n > 0) {
        const int bpv = (n + 63) / 64;
        hipLaunchKernelGGL(cullVisibilityViewsKernel, dim3(numViews * bpv), dim3(64), 0, s, views, in, n, bpv);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(cullCompactViewsKernel, dim3(numViews), dim3(kCullThreads), 0, s, views, in, n);
    return hipGetLastError();
}

hipError_t colorVpls(const sgi_dev::CubeArgs& diffuse, const sgi_dev::CubeArgs& dist, int distFormat,
                     const sgi_dev::FrameArgs& fr, float mult, const int32_t* visible, int maxCount,
                     const sgi_atlas_entry* entries, sgi_vpl* vpls, hipStream_t s) {
    if (maxCount <= 0) return hipSuccess;
    const f3 sun{fr.sunDir[0], fr.sunDir[1], fr.sunDir[2]};
    const f3 col{fr.sunColor[0], fr.sunColor[1], fr.sunColor[2]};
    const dim3 grid((unsigned)((maxCount + 255) / 256));
    if (distFormat == SGI_TEXEL_F16)
        hipLaunchKernelGGL(colorVplsKernel<SGI_TEXEL_F16>, grid, dim3(256), 0, s, diffuse, dist, sun, col, mult,
                           visible, maxCount, entries, vpls);
    else if (distFormat == SGI_TEXEL_UNORM16)
        hipLaunchKernelGGL(colorVplsKernel<SGI_TEXEL_UNORM16>, grid, dim3(256), 0, s, diffuse, dist, sun, col, mult,
                           visible, maxCount, entries, vpls);
    else
        hipLaunchKernelGGL(colorVplsKernel<SGI_TEXEL_F32>, grid, dim3(256), 0, s, diffuse, dist, sun, col, mult,
                           visible, maxCount, entries, vpls);
    return hipGetLastError();
}

hipError_t gatherVisibleEntriesViews(const sgi_batch::ViewDesc* views, int numViews, const sgi_atlas_entry* all,
                                     int maxCount, hipStream_t s) {
    if (maxCount <= 0) return hipSuccess;
    const int bpv = (maxCount + 255) / 256;
    hipLaunchKernelGGL(gatherEntriesViewsKernel, dim3(numViews * bpv), dim3(256), 0, s, views, all, maxCount, bpv);
    return hipGetLastError();
}

hipError_t cullTilesViews(const sgi_batch::ViewDesc* views, int numViews, int totalTiles, int tileSize, int K,
                          hipStream_t s) {
    switch (tileSize) {
        case 8:
            hipLaunchKernelGGL(cullTilesViewsKernel<8>, dim3(totalTiles), dim3(64), 0, s, views, numViews, totalTiles, K);
            break;
        case 16:
            hipLaunchKernelGGL(cullTilesViewsKernel<16>, dim3(totalTiles), dim3(64), 0, s, views, numViews, totalTiles,
                               K);
            break;
        case 32:
            hipLaunchKernelGGL(cullTilesViewsKernel<32>, dim3(totalTiles), dim3(64), 0, s, views, numViews, totalTiles,
                               K);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace sgi_launch
