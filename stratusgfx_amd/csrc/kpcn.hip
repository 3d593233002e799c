// Learned denoiser: the kernel-predicting encoder-decoder network's kernels.
//
// The convolutions are implicit GEMMs on MFMA (v_mfma_f32_16x16x32_bf16,
// BF16 operands, FP32 accumulation).  One MFMA multiplies a 16 x 32 block of
// weights (16 output channels x 32 input channels of one tap) by a 32 x 16
// block of the input (32 channels x 16 pixels of one image row), so the
// channels are the reduction dimension and each lane reads 8 consecutive
// channels (16 B) of one pixel.  The result lane layout puts 4 consecutive
// output channels of one pixel in each lane, which the epilogue stores as 8 B
// (BF16) or 16 B (FP32).
//
//   conv3x3Kernel  a 256-thread block computes 16x16 output pixels x 32
//                  channels: the 18x18 input halo is staged once in LDS, the
//                  wave's 4 rows x 2 channel blocks (8 accumulators) run the
//                  9 taps; the layer's weights (9 taps x 2 channel blocks,
//                  72 VGPRs) stay in registers while the block walks its
//                  contiguous run of tiles.
//   wgradKernel    weight gradient, reduction over pixels: the gradient tile
//                  and three column-shifted copies of the input halo are
//                  staged transposed (channel-major, 8 pixels per 16 B) so the
//                  pixel dimension becomes the MFMA reduction; per-block
//                  partial sums are reduced in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include "launch.h"
#include "../../include/sgi_kpcn.h"

namespace {

using namespace sgi_dev;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kC = SGI_KPCN_CHANNELS;
constexpr int kTS = 16;       // output tile edge
constexpr int kHS = kTS + 2;  // halo tile edge
constexpr int kW = SGI_KPCN_LAYER_WEIGHTS;
constexpr int kP = SGI_KPCN_LAYER_PARAMS;
constexpr float kRadianceScale = SGI_KPCN_RADIANCE_SCALE;

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
// FP32 -> BF16 rounded to the nearest even (gfx950's v_cvt_pk_bf16_f32; FP32
// denormals are kept, so every finite value rounds as the integer form
// (u + 0x7fff + ((u >> 16) & 1)) >> 16 does); f2bf2 packs two, lo first.
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ uint32_t f2bf2(float lo, float hi) {
    const bf16x2v v = {(__bf16)lo, (__bf16)hi};
    return __builtin_bit_cast(uint32_t, v);
}

// Zero the BF16 lanes of v whose mask lane is not > 0.
__device__ __forceinline__ uint4 gateBf16x8(uint4 v, uint4 m) {
    uint32_t vv[4] = {v.x, v.y, v.z, v.w};
    const uint32_t mm[4] = {m.x, m.y, m.z, m.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t lo = (int16_t)(mm[i] & 0xffffu) > 0 ? 0x0000ffffu : 0u;
        const uint32_t hi = (int16_t)(mm[i] >> 16) > 0 ? 0xffff0000u : 0u;
        vv[i] &= lo | hi;
    }
    return make_uint4(vv[0], vv[1], vv[2], vv[3]);
}

// The contiguous run of tiles of this block; blocks of one XCD get adjacent runs.
__device__ __forceinline__ void blockTiles(int nTiles, int& t0, int& t1) {
    const int L = (int)xcdSwizzle(blockIdx.x, gridDim.x);
    const int per = (nTiles + (int)gridDim.x - 1) / (int)gridDim.x;
    t0 = min(nTiles, L * per);
    t1 = min(nTiles, t0 + per);
}

__device__ __forceinline__ bf16x8 asBf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }
// A 16-byte LDS read as one vector load (a uint4 read is four scalar loads that
// the backend re-merges without the alias scopes its wait insertion uses to
// tell the halo from the blocks being filled by LDS DMA).
__device__ __forceinline__ bf16x8 ldsBf16x8(const uint4* p) { return *reinterpret_cast<const bf16x8*>(p); }

// ln(1 + x), x >= 0, through the hardware log2 (v_log_f32, 1 ulp): 1 + x
// rounds away at most 2^-24 of x, so the result is within 2^-24 absolute and
// a few 2^-24 relative of log1p(x); the features it feeds are stored as BF16
// (2^-9 relative).
constexpr float kLn2 = 0.693147180559945f;
__device__ __forceinline__ float log1pFeature(float x) { return __builtin_amdgcn_logf(1.0f + x) * kLn2; }

// Network input features of pixel i (the 16 packed BF16 pairs of its 32
// channels) and its noisy radiance reservoir.rgb / max(reservoir.a, 1e-5).
// The features are rounded to BF16 right away, so they are formed with the
// hardware reciprocal, square root and logarithm (each within a few FP32 ulp)
// rather than the correctly rounded forms the GI passes decide with; the FP32
// radiance the reconstruction filters is within 2 ulp of the quotient.
__device__ __forceinline__ void pixelFeatures(const GbufArgs& g, const FrameArgs& fr,
                                              const float4* __restrict__ reservoir, int x, int y, uint32_t* o,
                                              float* L) {
    const size_t i = (size_t)y * g.W + x;
    const float4 r = reservoir[i];
    const float inv = __builtin_amdgcn_rcpf(gmax(kEPS, r.w));
    L[0] = r.x * inv;
    L[1] = r.y * inv;
    L[2] = r.z * inv;
    // world position: invPV (x_ndc, y_ndc, 2 depth - 1, 1), divided by w
    float w4[4];
    mulMat4(fr.invPV, ((float)x + 0.5f) * (2.0f / (float)g.W) - 1.0f, ((float)y + 0.5f) * (2.0f / (float)g.H) - 1.0f,
            g.depth[i] * 2.0f - 1.0f, 1.0f, w4);
    const float iw = __builtin_amdgcn_rcpf(w4[3]);
    const f3 d = ld3(fr.viewPos) - f3{w4[0] * iw, w4[1] * iw, w4[2] * iw};
    const f3 nr = gNormalRaw(g, i) * 2.0f - splat(1.0f);
    const f3 n = nr * __builtin_amdgcn_rsqf(dot(nr, nr));
    float f[SGI_KPCN_FEATURES];
    f[0] = log1pFeature(gmax(L[0], 0.0f) * kRadianceScale);
    f[1] = log1pFeature(gmax(L[1], 0.0f) * kRadianceScale);
    f[2] = log1pFeature(gmax(L[2], 0.0f) * kRadianceScale);
    const float4 al = gAlbedo(g, i);
    f[3] = al.x;
    f[4] = al.y;
    f[5] = al.z;
    f[6] = n.x;
    f[7] = n.y;
    f[8] = n.z;
    f[9] = log1pFeature(__builtin_amdgcn_sqrtf(dot(d, d))) * 0.125f;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int a = 2 * j, b = 2 * j + 1;
        const uint32_t lo = a < SGI_KPCN_FEATURES ? f2bf(f[a]) : 0u;
        const uint32_t hi = b < SGI_KPCN_FEATURES ? f2bf(f[b]) : 0u;
        o[j] = lo | (hi << 16);
    }
}

// Network input features and the noisy radiance.  CH = 32: the [h][w][32]
// layout of every activation (training, sgi_kpcn_features); CH = 16: the
// inference form [h][w][16] (channels 16..31 are zero, so the first
// convolution stages zeros for them instead of reading them).
template <int CH>
__global__ void featuresKernel(GbufArgs g, FrameArgs fr, const float4* __restrict__ reservoir,
                               uint16_t* __restrict__ feat, float* __restrict__ radiance) {
    static_assert(CH == 16 || CH == 32, "feature planes hold 16 or 32 channels");
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.W * g.H) return;
    uint32_t o[16];
    float L[3];
    pixelFeatures(g, fr, reservoir, i % g.W, i / g.W, o, L);
    radiance[3 * i + 0] = L[0];
    radiance[3 * i + 1] = L[1];
    radiance[3 * i + 2] = L[2];
    uint4* dst = reinterpret_cast<uint4*>(feat + (size_t)i * CH);
    dst[0] = make_uint4(o[0], o[1], o[2], o[3]);
    dst[1] = make_uint4(o[4], o[5], o[6], o[7]);
    if constexpr (CH == 32) {
        dst[2] = make_uint4(o[8], o[9], o[10], o[11]);
        dst[3] = make_uint4(o[12], o[13], o[14], o[15]);
    }
}

constexpr float kLog2e = 1.44269504088896f;

// Softmax weights of the 25 filter logits (max-subtracted; exp through the
// hardware exp2 of lg log2(e) - m log2(e), one FMA per tap: every argument is
// at most the rounding error of m log2(e), so no weight exceeds 1 + 2^-23).
__device__ __forceinline__ void softmax25(const float* lg, float* wgt) {
    float m = lg[0];
#pragma unroll
    for (int j = 1; j < SGI_KPCN_TAPS; ++j) m = gmax(m, lg[j]);
    const float mScaled = m * kLog2e;
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < SGI_KPCN_TAPS; ++j) {
        wgt[j] = __builtin_amdgcn_exp2f(__builtin_fmaf(lg[j], kLog2e, -mScaled));
        s += wgt[j];
    }
    const float inv = 1.0f / s;
#pragma unroll
    for (int j = 0; j < SGI_KPCN_TAPS; ++j) wgt[j] *= inv;
}

__device__ __forceinline__ void loadLogits(const float* logits, size_t p, float* lg) {
    const float4* src = reinterpret_cast<const float4*>(logits + p * kC);
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        const float4 v = src[j];
        lg[4 * j] = v.x;
        lg[4 * j + 1] = v.y;
        lg[4 * j + 2] = v.z;
        lg[4 * j + 3] = v.w;
    }
}

// One filter tap: o + v w per channel, one rounding each (FMA).
__device__ __forceinline__ f3 fmaTap(f3 v, float w, f3 o) {
    return f3{__builtin_fmaf(v.x, w, o.x), __builtin_fmaf(v.y, w, o.y), __builtin_fmaf(v.z, w, o.z)};
}

__device__ __forceinline__ f3 filtered(const float* radiance, int W, int H, int x, int y, const float* wgt) {
    f3 o = splat(0.0f);
#pragma unroll
    for (int j = 0; j < SGI_KPCN_TAPS; ++j) {
        const int xx = min(max(x + j % 5 - 2, 0), W - 1), yy = min(max(y + j / 5 - 2, 0), H - 1);
        o = fmaTap(ld3(radiance + 3 * ((size_t)yy * W + xx)), wgt[j], o);
    }
    return o;
}

// One axis of 2x bilinear upsampling: fine pixel x samples the coarser grid
// (n pixels) at (x + 0.5) / 2 - 0.5, i.e. coarse pixels (x - 1) / 2 and
// (x + 1) / 2 with weights 1/4, 3/4 (x even) or 3/4, 1/4 (x odd), clamped.
__device__ __forceinline__ void bilinearAxis(int x, int n, int& i0, int& i1, float& w0, float& w1) {
    const int k = x >> 1;
    const bool odd = (x & 1) != 0;
    const int a = odd ? k : k - 1;
    w0 = odd ? 0.75f : 0.25f;
    w1 = 1.0f - w0;
    i0 = min(max(a, 0), n - 1);
    i1 = min(max(a + 1, 0), n - 1);
}

// up(coarse) at fine pixel (x, y); T.coarse(X, Y) reads the coarser level
// ((h, w) its size) at in-range coordinates.
template <class T>
__device__ __forceinline__ f3 upBilinearT(const T& src, int h, int w, int x, int y) {
    int y0, y1, x0, x1;
    float wy0, wy1, wx0, wx1;
    bilinearAxis(y, h, y0, y1, wy0, wy1);
    bilinearAxis(x, w, x0, x1, wx0, wx1);
    const f3 top = src.coarse(x0, y0) * wx0 + src.coarse(x1, y0) * wx1;
    const f3 bot = src.coarse(x0, y1) * wx0 + src.coarse(x1, y1) * wx1;
    return top * wy0 + bot * wy1;
}

__device__ __forceinline__ float sigmoidf(float x) {
    return 1.0f / (1.0f + __builtin_amdgcn_exp2f(gmin(-x * kLog2e, 126.0f)));
}

// One level of the reconstruction at pixel (x, y): the 5x5 softmax filter of
// this level's radiance, blended with the upsampled coarser output when given.
// T.radiance(xx, yy) reads the level's radiance at in-image coordinates.
template <class T>
__device__ __forceinline__ f3 composePixelT(const float* lg, const T& src, bool hasCoarse, int W, int H, int x,
                                            int y) {
    float wgt[SGI_KPCN_TAPS];
    softmax25(lg, wgt);
    f3 o = splat(0.0f);
#pragma unroll
    for (int j = 0; j < SGI_KPCN_TAPS; ++j) {
        const int xx = min(max(x + j % 5 - 2, 0), W - 1), yy = min(max(y + j / 5 - 2, 0), H - 1);
        o = fmaTap(src.radiance(xx, yy), wgt[j], o);
    }
    if (hasCoarse) {
        const float b = sigmoidf(lg[SGI_KPCN_BLEND_CHANNEL]);
        o = o + (upBilinearT(src, (H + 1) / 2, (W + 1) / 2, x, y) - o) * b;
    }
    return o;
}

// The same taps read from global memory.
struct GlobalTaps {
    const float* rad;
    const float* co;
    int W, cw;
    __device__ __forceinline__ f3 radiance(int xx, int yy) const { return ld3(rad + 3 * ((size_t)yy * W + xx)); }
    __device__ __forceinline__ f3 coarse(int X, int Y) const { return ld3(co + 3 * ((size_t)Y * cw + X)); }
};

__device__ __forceinline__ f3 upBilinear(const float* __restrict__ coarse, int h, int w, int x, int y) {
    const GlobalTaps t{nullptr, coarse, 0, w};
    return upBilinearT(t, h, w, x, y);
}

__device__ __forceinline__ f3 composePixel(const float* lg, const float* __restrict__ radiance,
                                           const float* __restrict__ coarse, int W, int H, int x, int y) {
    const GlobalTaps t{radiance, coarse, W, (W + 1) / 2};
    return composePixelT(lg, t, coarse != nullptr, W, H, x, y);
}

// Where the convolution's input comes from (staged into the LDS halo):
//   IN_PLAIN     a [h][w][32] (gated by mask when given);
//   IN_PLAIN16   a [h][w][16], channels 16..31 staged as zeros;
//   IN_UPADD     a + up(b): the skip tensor plus the coarser level (nearest);
//                both chunks are prefetched raw and added when stored to LDS,
//                so the loads stay in flight across the previous tile's MFMAs.
// and where the result goes:
//   OUT_BF16 / OUT_F32  [h][w][32] with bias (and ReLU for BF16 when relu != 0);
//   OUT_BF16_POOL       OUT_BF16 plus the 2x2 mean of the stored BF16 values
//                       ([ceil(h/2)][ceil(w/2)][32], the next level's input),
//                       formed in the epilogue: a wave's rows 4w .. 4w + 3
//                       hold two row pairs, and the column pair sits in
//                       neighbouring lanes;
//   OUT_APPLY           the 27 used logits stay in LDS and the epilogue runs
//                       one level of the reconstruction (composePixel: the
//                       5x5 softmax filter of the level's radiance, blended
//                       with the coarser output when given; out [h][w][3])
//                       and writes the error logit (err [h][w]).
// Each fused form computes exactly what the separate kernels compute.
enum { IN_PLAIN = 0, IN_UPADD = 1, IN_PLAIN16 = 2 };
enum { OUT_BF16 = 0, OUT_F32 = 1, OUT_APPLY = 2, OUT_BF16_POOL = 3 };

struct ConvSrc {
    const uint16_t* a;
    const uint16_t* b;
    const uint16_t* mask;
    uint16_t* pooled;       // OUT_BF16_POOL
    const float* radiance;  // OUT_APPLY: this level's radiance
    const float* coarse;    // OUT_APPLY: the coarser level's output, or NULL
    // level-0 head (IN_UPADD, OUT_APPLY) with the error output: when flags is set, the epilogue
    // also forms each tile's fallback flag (flags[tile]; flagCounts[1 + tile] when set; tile 0
    // resets dilatedList[0] and flagCounts[0]) against flagThreshold
    uint8_t* flags;
    int32_t* dilatedList;
    int32_t* flagCounts;
    float flagThreshold;
};

__device__ __forceinline__ uint32_t addBf16x2(uint32_t a, uint32_t b) {
    const float lo = bf2f((uint16_t)(a & 0xffffu)) + bf2f((uint16_t)(b & 0xffffu));
    const float hi = bf2f((uint16_t)(a >> 16)) + bf2f((uint16_t)(b >> 16));
    return f2bf2(lo, hi);
}

__device__ __forceinline__ uint4 addChunk(uint4 a, uint4 b) {
    return make_uint4(addBf16x2(a.x, b.x), addBf16x2(a.y, b.y), addBf16x2(a.z, b.z), addBf16x2(a.w, b.w));
}

// Activation formats of the convolution: BF16 (64 B per 32-channel pixel; the
// training and default inference format) and OCP FP8 E4M3 (32 B per pixel;
// per-channel power-of-two scales folded into the weights and into the
// epilogue's per-output-channel out = fma(acc, k1, k2)).
enum { ACT_BF16 = 0, ACT_F8 = 1 };

template <int ACT>
struct Act {
    static constexpr int kPixBytes = ACT == ACT_F8 ? 32 : 64;  // a 32-channel pixel
    static constexpr int kChunks = kPixBytes / 16;              // its 16-byte chunks
    static constexpr int kPxU4 = ACT == ACT_F8 ? 3 : 5;         // LDS uint4 per staged pixel, one pad (bank spread)
    // the halo chunks one thread stages (every load issued before any is used):
    // chunk i = 16 bytes (c = i % kChunks) of halo pixel i / kChunks
    static constexpr int kStageSlots = (kHS * kHS * kChunks + 255) / 256;
};

typedef float f32x2 __attribute__((ext_vector_type(2)));

// FP8 E4M3 <-> FP32 with gfx950's conversions (OCP E4M3, round to nearest even);
// values are clamped to the finite range +-448 first, so the conversion's
// out-of-range behaviour is never reached.
__device__ __forceinline__ uint32_t packF8x4(float a, float b, float c, float d) {
    constexpr float m = 448.0f;
    const uint32_t lo = __builtin_amdgcn_cvt_pk_fp8_f32(clampv(a, -m, m), clampv(b, -m, m), 0, false);
    return __builtin_amdgcn_cvt_pk_fp8_f32(clampv(c, -m, m), clampv(d, -m, m), lo, true);
}

__device__ __forceinline__ void unpackF8x4(uint32_t w, float* o) {
    const f32x2 lo = __builtin_amdgcn_cvt_pk_f32_fp8(w, false), hi = __builtin_amdgcn_cvt_pk_f32_fp8(w, true);
    o[0] = lo.x;
    o[1] = lo.y;
    o[2] = hi.x;
    o[3] = hi.y;
}

// e + up(d) of FP8 activations sharing one scale: the sum of two E4M3 values is
// exact in FP32, so the only rounding is the store's.
__device__ __forceinline__ uint32_t addF8x4(uint32_t a, uint32_t b) {
    float x[4], y[4];
    unpackF8x4(a, x);
    unpackF8x4(b, y);
    return packF8x4(x[0] + y[0], x[1] + y[1], x[2] + y[2], x[3] + y[3]);
}

template <int ACT>
__device__ __forceinline__ uint4 addChunkT(uint4 a, uint4 b) {
    if constexpr (ACT == ACT_F8)
        return make_uint4(addF8x4(a.x, b.x), addF8x4(a.y, b.y), addF8x4(a.z, b.z), addF8x4(a.w, b.w));
    else
        return addChunk(a, b);
}

// IN_UPADD: the coarser operand is kept raw in vb and added at the LDS store
// (adding it here would wait for both loads before the MFMAs; the OUT_APPLY
// form loads only the skip operand here and stages the coarser one through
// LDS: stageUp). MASKED: gate IN_PLAIN by src.mask when it is given.
template <int ACT, int IN, bool DEFER, bool MASKED = IN == IN_PLAIN>
__device__ __forceinline__ void loadHalo(const ConvSrc& src, int H, int W, int tile, int tilesX, uint4* v,
                                         uint4* vb) {
    using A = Act<ACT>;
    static_assert(IN != IN_UPADD || DEFER, "the up-add operand is added at the LDS store");
    static_assert(!MASKED || ACT == ACT_BF16, "the gradient gate is a BF16 (training) form");
    constexpr int kInBytes = IN == IN_PLAIN16 ? A::kPixBytes / 2 : A::kPixBytes;
    const char* a = reinterpret_cast<const char*>(src.a);
    const char* b = reinterpret_cast<const char*>(src.b);
    const int x0 = (tile % tilesX) * kTS - 1, y0 = (tile / tilesX) * kTS - 1;
#pragma unroll
    for (int k = 0; k < A::kStageSlots; ++k) {
        const int i = (int)threadIdx.x + k * 256;
        const int px = i / A::kChunks, c = i % A::kChunks;
        const int gy = y0 + px / kHS, gx = x0 + px % kHS;
        v[k] = make_uint4(0u, 0u, 0u, 0u);
        if constexpr (IN == IN_UPADD) vb[k] = make_uint4(0u, 0u, 0u, 0u);
        if (IN == IN_PLAIN16 && c >= A::kChunks / 2) continue;
        if (i < kHS * kHS * A::kChunks && gx >= 0 && gx < W && gy >= 0 && gy < H) {
            const size_t g = ((size_t)gy * W + gx) * kInBytes + 16 * c;
            v[k] = *reinterpret_cast<const uint4*>(a + g);
            if constexpr (IN == IN_UPADD) {
                const int w2 = (W + 1) / 2;
                vb[k] = *reinterpret_cast<const uint4*>(b + ((size_t)(gy / 2) * w2 + gx / 2) * A::kPixBytes + 16 * c);
            } else if (MASKED && src.mask) {
                v[k] = gateBf16x8(v[k], *reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(src.mask) + g));
            }
        }
    }
}

// LDS DMA (global_load_lds) of 12 or 16 bytes per lane: lane l of the wave
// writes bytes 16 l .. 16 l + 11 (12-byte form) or .. 16 l + 15 (16-byte
// form) past lds (checked on gfx950). These forms exist on gfx950 only, so
// the builtin is checked and emitted in the device pass alone (the host pass
// would reject it and drop the kernel). lds must be a pointer expression into
// a __shared__ array.
__device__ __forceinline__ void dmaToLds12(const void* g, void* lds) {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_global_load_lds(g, lds, 12, 0, 0);
#endif
}

__device__ __forceinline__ void dmaToLds16(const void* g, void* lds) {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_global_load_lds(g, lds, 16, 0, 0);
#endif
}

// IN_UPADD's coarser operand for one tile, staged in LDS by DMA so that no
// registers hold it across the MFMAs: halo rows ty*kTS - 1 .. ty*kTS + kTS of
// the finer level read coarser rows ty*kTS/2 - 1 .. ty*kTS/2 + kTS/2, a kUpHS
// x kUpHS block at (tx*kTS/2 - 1, ty*kTS/2 - 1) whose 16-byte chunk d (pixel
// d / kChunks, chunk d % kChunks) sits at index d, read clamped to the coarser
// image (every entry an in-image halo pixel reads is in range).
constexpr int kUpHS = kTS / 2 + 2;

template <int ACT>
struct UpStage {
    static constexpr int kChunks = kUpHS * kUpHS * Act<ACT>::kChunks;
    static constexpr int kSlots = (kChunks + 255) / 256;
};

template <int ACT>
__device__ __forceinline__ void stageUp(const ConvSrc& src, int H, int W, int tile, int tilesX, uint4* upS) {
    using A = Act<ACT>;
    const char* b = reinterpret_cast<const char*>(src.b);
    const int ch = (H + 1) / 2, cw = (W + 1) / 2;
    const int cx0 = (tile % tilesX) * (kTS / 2) - 1, cy0 = (tile / tilesX) * (kTS / 2) - 1;
    const int wave = (int)threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < UpStage<ACT>::kSlots; ++k) {
        const int d = (int)threadIdx.x + k * 256;
        size_t g = 0;
        if (d < UpStage<ACT>::kChunks) {
            const int px = d / A::kChunks, c = d % A::kChunks;
            const int gy = min(max(cy0 + px / kUpHS, 0), ch - 1), gx = min(max(cx0 + px % kUpHS, 0), cw - 1);
            g = ((size_t)gy * cw + gx) * A::kPixBytes + 16 * c;
        }
        dmaToLds16(b + g, upS + k * 256 + wave * 64);
    }
}

// Floats per pixel of the staged logits: channels 0..26 used, 27 padding, so a
// lane's 4 channels move as one 16-byte LDS access (28 dwords per pixel puts the
// 16 pixels of a 16-lane group on disjoint banks).
constexpr int kLogitStride = 28;
constexpr int kRadHS = kTS + 4;     // radiance staged for the 5x5 taps (OUT_APPLY)
constexpr int kCoHS = kTS / 2 + 2;  // the coarser output staged for up() (OUT_APPLY)
constexpr int kRadSlots = (kRadHS * kRadHS + 255) / 256;  // staged pixels per thread
constexpr int kCoSlots = (kCoHS * kCoHS + 255) / 256;

// OUT_APPLY's blocks of one tile, loaded into registers one tile ahead and
// stored to LDS at the tile's start: its radiance (2-pixel margin, kRadHS^2
// pixels) and its coarser output (1-pixel margin, kCoHS^2 pixels), clamped to
// the image as the global taps clamp. Two pixels per thread: thread t holds
// radiance pixel t, and radiance pixel t + 256 or, past the radiance block's
// end, coarser pixel t + 256 - kRadHS^2 (the coarser block fills threads the
// radiance block leaves idle). Loads are never skipped: a thread past both
// blocks, or a coarser pixel with no coarser level, loads an in-range pixel it
// does not store.
static_assert(kRadSlots == 2 && kRadHS * kRadHS + kCoHS * kCoHS <= 512, "two staged pixels per thread");
struct ApplyRegs {
    float v[2][3];
};

__device__ __forceinline__ void loadApply(const float* radiance, const float* coarse, int H, int W, int tile,
                                          int tilesX, ApplyRegs& r) {
    const int tx = tile % tilesX, ty = tile / tilesX;
    const int rx0 = tx * kTS - 2, ry0 = ty * kTS - 2;
    const int cx0 = tx * (kTS / 2) - 1, cy0 = ty * (kTS / 2) - 1;
    const int ch = (H + 1) / 2, cw = (W + 1) / 2;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int i = (int)threadIdx.x + k * 256;
        const float* q;
        if (i < kRadHS * kRadHS || !coarse) {
            const int j = min(i, kRadHS * kRadHS - 1);
            const int gy = min(max(ry0 + j / kRadHS, 0), H - 1), gx = min(max(rx0 + j % kRadHS, 0), W - 1);
            q = radiance + 3 * ((size_t)gy * W + gx);
        } else {
            const int j = min(i - kRadHS * kRadHS, kCoHS * kCoHS - 1);
            const int gy = min(max(cy0 + j / kCoHS, 0), ch - 1), gx = min(max(cx0 + j % kCoHS, 0), cw - 1);
            q = coarse + 3 * ((size_t)gy * cw + gx);
        }
        r.v[k][0] = q[0];
        r.v[k][1] = q[1];
        r.v[k][2] = q[2];
    }
}

__device__ __forceinline__ void storeApply(const ApplyRegs& r, bool hasCoarse, f32x4* radS, f32x4* coS) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int i = (int)threadIdx.x + k * 256;
        const f32x4 v = f32x4{r.v[k][0], r.v[k][1], r.v[k][2], 0.0f};
        if (i < kRadHS * kRadHS)
            radS[i] = v;
        else if (hasCoarse && i < kRadHS * kRadHS + kCoHS * kCoHS)
            coS[i - kRadHS * kRadHS] = v;
    }
}

// One level of the reconstruction at tile pixel (lx, ly) from the LDS-staged
// blocks (one pixel per 16 bytes, rgb in the first three floats): radiance
// staged with a 2-pixel margin and the coarser output with a 1-pixel margin,
// both read clamped to their images when staged, so every tap the clamped
// formula names sits at a constant offset from the pixel's base: the same
// values, summed in composePixelT's order, with one immediate-offset LDS read
// per tap and no per-tap index arithmetic.
__device__ __forceinline__ f3 composePixelStaged(const float* lg, const f32x4* rad, const f32x4* co, bool hasCoarse,
                                                 int lx, int ly) {
    float wgt[SGI_KPCN_TAPS];
    softmax25(lg, wgt);
    const f32x4* r = rad + ly * kRadHS + lx;  // tap (dx, dy) of the pixel: r[(dy + 2) kRadHS + dx + 2]
    f3 o = splat(0.0f);
#pragma unroll
    for (int j = 0; j < SGI_KPCN_TAPS; ++j) {
        const f32x4 v = r[(j / 5) * kRadHS + j % 5];
        o = fmaTap(f3{v.x, v.y, v.z}, wgt[j], o);
    }
    if (hasCoarse) {
        // bilinearAxis: coarse pixels a, a + 1 (a = x / 2 for odd x, x / 2 - 1 for even), clamped;
        // in the staged block (first pixel tx*kTS/2 - 1) a sits at (lx + 1) / 2
        const int ax = (lx + 1) >> 1, ay = (ly + 1) >> 1;
        const float wx0 = (lx & 1) ? 0.75f : 0.25f, wy0 = (ly & 1) ? 0.75f : 0.25f;
        const float wx1 = 1.0f - wx0, wy1 = 1.0f - wy0;
        auto c = [&](int X, int Y) {
            const f32x4 v = co[Y * kCoHS + X];
            return f3{v.x, v.y, v.z};
        };
        const f3 top = c(ax, ay) * wx0 + c(ax + 1, ay) * wx1;
        const f3 bot = c(ax, ay + 1) * wx0 + c(ax + 1, ay + 1) * wx1;
        const float b = sigmoidf(lg[SGI_KPCN_BLEND_CHANNEL]);
        o = o + (top * wy0 + bot * wy1 - o) * b;
    }
    return o;
}

// The MFMA operand type of an activation format: 8 BF16 or 8 FP8 values.
template <int ACT>
struct MfmaOp {
    typedef bf16x8 T;
};
template <>
struct MfmaOp<ACT_F8> {
    typedef long T;
};

// AIN: the input's format (halo, weights and MFMA); AOUT: the format an activation
// output (OUT_BF16 / OUT_BF16_POOL) is stored in. When either is FP8 the epilogue
// is scaled: bias holds [k1[32], k2[32]] and a channel's value is fma(acc, k1, k2)
// (otherwise acc + bias).
// Blocks resident per CU: two (VGPRs near 256, and for OUT_APPLY the staged blocks in LDS),
// three for FP8 inputs on the plain path (FP8 weights take 36 VGPRs instead of 72 and the
// FP8 halo 15.5 KB of LDS; the deferred up-add's raw chunks would not fit the 168 VGPRs)
// and for the first convolution with its weights in LDS.
// The first convolution (BF16, 16-channel features in) keeps its weight fragments in
// LDS, one copy per block (9 taps x 2 channel blocks x 64 lanes x 16 B = 18 KB), each
// wave reading a tap's two fragments before its MFMAs: without the 72 VGPRs they take in
// registers it runs three blocks per CU (75 -> 72 us at 1080p). On the other layers the
// extra LDS round trip before the first MFMA cost more than the third block gave.
template <int IN, int AIN>
constexpr bool weightsInLds() {
    return AIN == ACT_BF16 && IN == IN_PLAIN16;
}

template <int IN, int OUT, int AIN>
constexpr int convBlocksPerCu() {
    return OUT != OUT_APPLY && ((AIN == ACT_F8 && IN == IN_PLAIN) || weightsInLds<IN, AIN>()) ? 3 : 2;
}

template <int IN, int OUT, int AIN, int AOUT>
__global__ __launch_bounds__(256, (convBlocksPerCu<IN, OUT, AIN>())) void conv3x3Kernel(
    ConvSrc src, const void* __restrict__ w, const float* __restrict__ bias, int relu, int H, int W, int tilesX,
    int nTiles, void* __restrict__ out, float* __restrict__ err) {
    using A = Act<AIN>;
    constexpr bool kScaled = AIN == ACT_F8 || AOUT == ACT_F8;
    constexpr int kHaloU4 = kHS * kHS * A::kPxU4;
    constexpr int kLogitU4 = kTS * kTS * kLogitStride / 4;
    constexpr int kSmemU4 = OUT == OUT_APPLY && kLogitU4 > kHaloU4 ? kLogitU4 : kHaloU4;
    __shared__ uint4 smem[kSmemU4];  // the input halo, then (OUT_APPLY) the tile's logits
    __shared__ f32x4 radS[OUT == OUT_APPLY ? kRadSlots * 256 : 1];
    __shared__ f32x4 coS[OUT == OUT_APPLY ? kCoSlots * 256 : 1];
    // IN_UPADD with OUT_APPLY: the coarser operand staged by stageUp (the
    // epilogue needs the registers a deferred add would hold)
    constexpr bool kLdsUp = IN == IN_UPADD && OUT == OUT_APPLY;
    constexpr bool kDefer = IN == IN_UPADD && !kLdsUp;
    constexpr int kHaloIn = kLdsUp ? IN_PLAIN : IN;  // the skip operand loads as a plain input
    __shared__ uint4 upS[IN == IN_UPADD && OUT == OUT_APPLY ? UpStage<AIN>::kSlots * 256 : 1];
    constexpr bool kWLds = weightsInLds<IN, AIN>();
    __shared__ uint4 wS[kWLds ? 9 * 2 * 64 : 1];  // fragment (t, mb) of lane l at (2 t + mb) 64 + l
    constexpr bool kFlags = IN == IN_UPADD && OUT == OUT_APPLY;
    __shared__ float flagS[kFlags ? 256 : 1];  // the tile's exp(min(err, 80)), pixel by pixel
    uint4* halo = smem;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lg = lane >> 4;
    // A operands: W[t][mb*16 + l16][8 lg .. 8 lg + 7]
    typename MfmaOp<AIN>::T wa[kWLds ? 1 : 9][2];
    if constexpr (kWLds) {
        // visible to every wave after the first tile's barriers, before its first MFMA
        for (int i = (int)threadIdx.x; i < 9 * 2 * 64; i += 256) {
            const int t = i >> 7, mb = (i >> 6) & 1, ln = i & 63;
            const size_t e = (size_t)(t * kC + mb * 16 + (ln & 15)) * kC + 8 * (ln >> 4);
            wS[i] = *reinterpret_cast<const uint4*>(static_cast<const uint16_t*>(w) + e);
        }
    } else {
#pragma unroll
        for (int t = 0; t < 9; ++t)
#pragma unroll
            for (int mb = 0; mb < 2; ++mb) {
                const size_t e = (size_t)(t * kC + mb * 16 + l16) * kC + 8 * lg;
                if constexpr (AIN == ACT_F8)
                    wa[t][mb] = *reinterpret_cast<const long*>(static_cast<const uint8_t*>(w) + e);
                else
                    wa[t][mb] = asBf16x8(*reinterpret_cast<const uint4*>(static_cast<const uint16_t*>(w) + e));
            }
    }
    float k1v[2][4], bv[2][4];  // scaled: (k1, k2); otherwise bv = bias
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int o = mb * 16 + 4 * lg + r;
            k1v[mb][r] = kScaled ? bias[o] : 1.0f;
            bv[mb][r] = kScaled ? bias[kC + o] : (bias ? bias[o] : 0.0f);
        }

    int t0, t1;
    blockTiles(nTiles, t0, t1);
    uint4 pre[A::kStageSlots], preB[kDefer ? A::kStageSlots : 1];
    ApplyRegs apl;  // OUT_APPLY: the staged blocks' values, one tile ahead
    if (t0 < t1) {
        loadHalo<AIN, kHaloIn, kDefer, !kLdsUp && IN == IN_PLAIN && AIN == ACT_BF16>(src, H, W, t0, tilesX, pre,
                                                                                     preB);
        if constexpr (kLdsUp) stageUp<AIN>(src, H, W, t0, tilesX, upS);
        if constexpr (OUT == OUT_APPLY) loadApply(src.radiance, src.coarse, H, W, t0, tilesX, apl);
    }
    // the fallback flag of a finished tile: the mean of exp(min(err, 80)) over its in-image
    // pixels (terms in flagS, 0 outside the image) against the threshold (every tile when the
    // threshold is <= 0), summed as a fixed binary tree over the 256 pixels (pixel i + 128 into i,
    // then i + 64 into i, ... 1): lane l of wave 0 adds pixels l, l + 64, l + 128, l + 192 in
    // registers and the last six levels through shuffles; tile 0 also resets the counters
    // dilateListKernel accumulates into
    int prevTile = -1;
    auto flagTile = [&](int ft) {
        if (wave != 0) return;
        float sum = (flagS[lane] + flagS[lane + 128]) + (flagS[lane + 64] + flagS[lane + 192]);
#pragma unroll
        for (int sft = 32; sft > 0; sft >>= 1) sum = sum + __shfl_down(sum, sft);
        if (lane == 0) {
            const int fx = ft % tilesX, fy = ft / tilesX;
            const int n = (min(W, fx * kTS + kTS) - fx * kTS) * (min(H, fy * kTS + kTS) - fy * kTS);
            const float thr = src.flagThreshold;
            const uint8_t f = (!(thr > 0.0f) || sum / (float)n > thr) ? 1 : 0;
            src.flags[ft] = f;
            if (src.flagCounts) src.flagCounts[1 + ft] = f;
            if (ft == 0) {
                src.dilatedList[0] = 0;
                if (src.flagCounts) src.flagCounts[0] = 0;
            }
        }
    };
    for (int tile = t0; tile < t1; ++tile) {
        const int tx = tile % tilesX, ty = tile / tilesX;
        // kLdsUp: this thread's staging DMA has landed before the barrier
        if constexpr (kLdsUp) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        // this tile's chunks were loaded into `pre` before the previous tile's MFMAs
#pragma unroll
        for (int k = 0; k < A::kStageSlots; ++k) {
            const int i = (int)threadIdx.x + k * 256;
            if (i < kHS * kHS * A::kChunks) {
                const int px = i / A::kChunks, c = i % A::kChunks;
                if constexpr (kDefer) {
                    halo[px * A::kPxU4 + c] = addChunkT<AIN>(pre[k], preB[k]);
                } else if constexpr (kLdsUp) {
                    // in-image halo pixels add their coarser pixel; the zero padding stays zero
                    const int gy = ty * kTS - 1 + px / kHS, gx = tx * kTS - 1 + px % kHS;
                    uint4 v = pre[k];
                    if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
                        const int q = ((gy >> 1) - (ty * (kTS / 2) - 1)) * kUpHS + (gx >> 1) - (tx * (kTS / 2) - 1);
                        v = addChunkT<AIN>(v, upS[q * A::kChunks + c]);
                    }
                    halo[px * A::kPxU4 + c] = v;
                } else {
                    halo[px * A::kPxU4 + c] = pre[k];
                }
            }
        }
        if constexpr (OUT == OUT_APPLY) storeApply(apl, src.coarse != nullptr, radS, coS);
        // the previous tile's terms were stored before this tile's first barrier and are next
        // written after its second
        if constexpr (kFlags)
            if (src.flags && prevTile >= 0) flagTile(prevTile);
        __syncthreads();
        if (tile + 1 < t1) {
            loadHalo<AIN, kHaloIn, kDefer, !kLdsUp && IN == IN_PLAIN && AIN == ACT_BF16>(src, H, W, tile + 1, tilesX,
                                                                                         pre, preB);
            if constexpr (kLdsUp) stageUp<AIN>(src, H, W, tile + 1, tilesX, upS);
            if constexpr (OUT == OUT_APPLY) loadApply(src.radiance, src.coarse, H, W, tile + 1, tilesX, apl);
        }
        // the bias is added to the finished sums (sum + bias, the oracle's order: starting the
        // sums from the bias moves FP32 results by an ulp, which decides BF16 rounding ties
        // differently, and the coarse levels' pooled values sit on ties often)
        f32x4 acc[2][4];
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) acc[mb][nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int dy = t / 3, dx = t % 3;
            typename MfmaOp<AIN>::T a0, a1;
            if constexpr (kWLds) {
                a0 = ldsBf16x8(wS + (2 * t) * 64 + lane);
                a1 = ldsBf16x8(wS + (2 * t + 1) * 64 + lane);
            } else {
                a0 = wa[t][0];
                a1 = wa[t][1];
            }
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                const int row = wave * 4 + nb;
                const int hp = (row + dy) * kHS + l16 + dx;  // the staged halo pixel
                if constexpr (AIN == ACT_F8) {
                    // 8 FP8 channels 8 lg .. 8 lg + 7: one 8-byte LDS read
                    const long b = *reinterpret_cast<const long*>(reinterpret_cast<const char*>(halo + hp * A::kPxU4) +
                                                                  8 * lg);
                    acc[0][nb] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a0, b, acc[0][nb], 0, 0, 0);
                    acc[1][nb] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a1, b, acc[1][nb], 0, 0, 0);
                } else {
                    const bf16x8 b = ldsBf16x8(halo + hp * A::kPxU4 + lg);
                    acc[0][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b, acc[0][nb], 0, 0, 0);
                    acc[1][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b, acc[1][nb], 0, 0, 0);
                }
            }
        }
        // lane holds out channels mb*16 + 4 lg + r of pixel (row, l16)
        if constexpr (OUT == OUT_APPLY) {
            float* logits = reinterpret_cast<float*>(smem);
            // every wave has read the halo the logits overwrite: its LDS reads have returned
            // (lgkmcnt) and no memory access moves across (the clobber). Not __syncthreads(),
            // whose fence would also wait for the LDS DMA and the next tile's prefetch in flight.
            asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
#pragma unroll
            for (int nb = 0; nb < 4; ++nb)
#pragma unroll
                for (int mb = 0; mb < 2; ++mb) {
                    const int co = mb * 16 + 4 * lg;  // this lane's 4 channels co .. co + 3
                    if (co < kLogitStride) {
                        f32x4 v;
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            v[r] = kScaled ? __builtin_fmaf(acc[mb][nb][r], k1v[mb][r], bv[mb][r])
                                           : acc[mb][nb][r] + bv[mb][r];
                        *reinterpret_cast<f32x4*>(logits + ((wave * 4 + nb) * kTS + l16) * kLogitStride + co) = v;
                    }
                }
            // the logits are in LDS before any thread reads them (the staged blocks were stored
            // before the tile's second barrier)
            asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
            const int x = tx * kTS + (int)threadIdx.x % kTS, y = ty * kTS + (int)threadIdx.x / kTS;
            float errV = 0.0f;
            if (x < W && y < H) {
                float lgv[kLogitStride];
#pragma unroll
                for (int j = 0; j < kLogitStride / 4; ++j) {
                    const f32x4 v = *reinterpret_cast<const f32x4*>(logits + threadIdx.x * kLogitStride + 4 * j);
                    lgv[4 * j] = v.x;
                    lgv[4 * j + 1] = v.y;
                    lgv[4 * j + 2] = v.z;
                    lgv[4 * j + 3] = v.w;
                }
                const int lx = (int)threadIdx.x % kTS, ly = (int)threadIdx.x / kTS;
                const f3 o = composePixelStaged(lgv, radS, coS, src.coarse != nullptr, lx, ly);
                const size_t p = (size_t)y * W + x;
                float* dst = static_cast<float*>(out) + 3 * p;
                dst[0] = o.x;
                dst[1] = o.y;
                dst[2] = o.z;
                if (err) err[p] = lgv[SGI_KPCN_ERR_CHANNEL];
                errV = lgv[SGI_KPCN_ERR_CHANNEL];
            }
            if constexpr (kFlags) {
                // this pixel's term of the tile's fallback flag, 0 outside the image (reduced by
                // wave 0 at the next tile's start, or after the loop)
                if (src.flags) flagS[threadIdx.x] = x < W && y < H ? expf(gmin(errV, 80.0f)) : 0.0f;
                prevTile = tile;
            }
        } else {
            constexpr bool kBf16 = OUT == OUT_BF16 || OUT == OUT_BF16_POOL;
            const int x = tx * kTS + l16;
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                const int y = ty * kTS + wave * 4 + nb;
                const size_t p = (size_t)y * W + x;
#pragma unroll
                for (int mb = 0; mb < 2; ++mb) {
                    float v[4];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        v[r] = kScaled ? __builtin_fmaf(acc[mb][nb][r], k1v[mb][r], bv[mb][r])
                                       : acc[mb][nb][r] + bv[mb][r];
                        if (kBf16 && relu) v[r] = gmax(v[r], 0.0f);
                    }
                    const int co = mb * 16 + 4 * lg;
                    if constexpr (OUT == OUT_F32) {
                        static_assert(!kScaled, "FP32 activations are a BF16 (training) form");
                        if (x < W && y < H)
                            *reinterpret_cast<float4*>(static_cast<float*>(out) + p * kC + co) =
                                make_float4(v[0], v[1], v[2], v[3]);
                    } else if constexpr (AOUT == ACT_F8) {
                        const uint32_t q = packF8x4(v[0], v[1], v[2], v[3]);
                        if (x < W && y < H) *reinterpret_cast<uint32_t*>(static_cast<uint8_t*>(out) + p * kC + co) = q;
                        if constexpr (OUT == OUT_BF16_POOL) {  // keep the stored values for the pooling
                            float d[4];
                            unpackF8x4(q, d);
#pragma unroll
                            for (int r = 0; r < 4; ++r) acc[mb][nb][r] = d[r];
                        }
                    } else {
                        const uint32_t b01 = f2bf2(v[0], v[1]), b23 = f2bf2(v[2], v[3]);
                        if (x < W && y < H)
                            *reinterpret_cast<uint2*>(static_cast<uint16_t*>(out) + p * kC + co) = make_uint2(b01, b23);
                        if constexpr (OUT == OUT_BF16_POOL) {  // keep the stored values for the pooling
                            acc[mb][nb][0] = bf2f((uint16_t)(b01 & 0xffffu));
                            acc[mb][nb][1] = bf2f((uint16_t)(b01 >> 16));
                            acc[mb][nb][2] = bf2f((uint16_t)(b23 & 0xffffu));
                            acc[mb][nb][3] = bf2f((uint16_t)(b23 >> 16));
                        }
                    }
                }
            }
            if constexpr (OUT == OUT_BF16_POOL) {
                // rows 4 wave + {0, 1} and {2, 3} pool into two coarse rows; the even lane
                // of each column pair sums its children in poolKernel's order
                const int cw = (W + 1) / 2;
                const bool xIn = x < W, x1In = x + 1 < W, even = (l16 & 1) == 0;
#pragma unroll
                for (int pr = 0; pr < 2; ++pr) {
                    const int y0 = ty * kTS + wave * 4 + 2 * pr;
                    const bool y1In = y0 + 1 < H;
                    const int n = (1 + (x1In ? 1 : 0)) * (1 + (y1In ? 1 : 0));
                    const float inv = n == 4 ? 0.25f : (n == 2 ? 0.5f : 1.0f);  // 1 / n, exact
#pragma unroll
                    for (int mb = 0; mb < 2; ++mb) {
                        float m[4];
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const float a00 = acc[mb][2 * pr][r], a10 = acc[mb][2 * pr + 1][r];
                            const float a01 = __shfl_xor(a00, 1), a11 = __shfl_xor(a10, 1);
                            float sum = 0.0f;
                            sum += a00;
                            if (x1In) sum += a01;
                            if (y1In) {
                                sum += a10;
                                if (x1In) sum += a11;
                            }
                            m[r] = sum * inv;
                        }
                        const size_t q = ((size_t)(y0 / 2) * cw + x / 2) * kC + mb * 16 + 4 * lg;
                        if constexpr (AOUT == ACT_F8) {
                            const uint32_t o = packF8x4(m[0], m[1], m[2], m[3]);
                            if (even && xIn && y0 < H) *reinterpret_cast<uint32_t*>(reinterpret_cast<uint8_t*>(src.pooled) + q) = o;
                        } else {
                            const uint32_t o01 = f2bf2(m[0], m[1]), o23 = f2bf2(m[2], m[3]);
                            if (even && xIn && y0 < H) *reinterpret_cast<uint2*>(src.pooled + q) = make_uint2(o01, o23);
                        }
                    }
                }
            }
        }
    }
    if constexpr (kFlags) {
        if (src.flags && prevTile >= 0) {  // the block's last tile
            __syncthreads();
            flagTile(prevTile);
        }
    }
}

// E4M3 bytes of n floats (clamped to +-448, nearest even): thread i converts
// values 4i .. 4i + 3 (those past n are left alone).
// BF16 of n floats (f2bf: nearest even).
__global__ __launch_bounds__(256) void encodeBf16Kernel(const float* __restrict__ in, int64_t n,
                                                        uint16_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = f2bf(in[i]);
}

__global__ __launch_bounds__(256) void encodeF8Kernel(const float* __restrict__ in, int64_t n, uint8_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x, j = 4 * i;
    if (j >= n) return;
    float v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = j + k < n ? in[j + k] : 0.0f;
    const uint32_t q = packF8x4(v[0], v[1], v[2], v[3]);
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (j + k < n) out[j + k] = (uint8_t)(q >> (8 * k));
}

// Weight gradient.  LDS: xs[s][ci][row][chunk] = 8 pixels x[y0 - 1 + row][x0 + 8 chunk + s - 1 .. +7][ci]
// (s = dx shift 0..2, row 0..17) and dz[co][row][chunk] = 8 pixels of the gated gradient.
__global__ __launch_bounds__(256) void wgradKernel(const uint16_t* __restrict__ x, const uint16_t* __restrict__ dy,
                                                   const uint16_t* __restrict__ mask, int H, int W, int tilesX,
                                                   int nTiles, float* __restrict__ partials) {
    __shared__ __attribute__((aligned(16))) uint16_t xs[3][kC][kHS][kTS];
    __shared__ __attribute__((aligned(16))) uint16_t dz[kC][kTS][kTS];
    __shared__ float dbPart[4][kC];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lg = lane >> 4;
    f32x4 acc[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float dbAcc = 0.0f;
    int t0, t1;
    blockTiles(nTiles, t0, t1);
    for (int tile = t0; tile < t1; ++tile) {
        const int tx = tile % tilesX, ty = tile / tilesX;
        const int bx = tx * kTS, by = ty * kTS;
        __syncthreads();
        // gated gradient tile, transposed
        for (int i = threadIdx.x; i < kTS * kTS * 4; i += 256) {
            const int px = i >> 2, c = i & 3;
            const int gy = by + px / kTS, gx = bx + px % kTS;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (gx < W && gy < H) {
                const size_t g = ((size_t)gy * W + gx) * kC + c * 8;
                v = *reinterpret_cast<const uint4*>(dy + g);
                if (mask) v = gateBf16x8(v, *reinterpret_cast<const uint4*>(mask + g));
            }
            const uint32_t vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                dz[c * 8 + 2 * j][px / kTS][px % kTS] = (uint16_t)(vv[j] & 0xffffu);
                dz[c * 8 + 2 * j + 1][px / kTS][px % kTS] = (uint16_t)(vv[j] >> 16);
            }
        }
        // input halo, transposed, in three column shifts
        for (int i = threadIdx.x; i < kHS * kHS * 4; i += 256) {
            const int px = i >> 2, c = i & 3;
            const int hy = px / kHS, hx = px % kHS;
            const int gy = by - 1 + hy, gx = bx - 1 + hx;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (gx >= 0 && gx < W && gy >= 0 && gy < H)
                v = *reinterpret_cast<const uint4*>(x + ((size_t)gy * W + gx) * kC + c * 8);
            const uint32_t vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                const int col = hx - s;
                if (col < 0 || col >= kTS) continue;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    xs[s][c * 8 + 2 * j][hy][col] = (uint16_t)(vv[j] & 0xffffu);
                    xs[s][c * 8 + 2 * j + 1][hy][col] = (uint16_t)(vv[j] >> 16);
                }
            }
        }
        __syncthreads();
        // bias gradient: lane pairs per channel
        if (threadIdx.x < 2 * kC) {
            const int co = threadIdx.x >> 1, half = threadIdx.x & 1;
            float s = 0.0f;
            for (int r = half * 8; r < half * 8 + 8; ++r)
                for (int c = 0; c < kTS; ++c) s += bf2f(dz[co][r][c]);
            dbAcc += s;
        }
        // 8 reduction steps of 32 pixels (two tile rows); combo q = wave * 9 + i:
        // tap = q / 4, mb = (q / 2) % 2, nb = q % 2
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int prow = 2 * k + (lg >> 1), chunk = lg & 1;
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                const int q = wave * 9 + i;
                const int tap = q >> 2, mb = (q >> 1) & 1, nb = q & 1;
                const int tdy = tap / 3, tdx = tap % 3;
                const bf16x8 a = asBf16x8(*reinterpret_cast<const uint4*>(&dz[mb * 16 + l16][prow][chunk * 8]));
                const bf16x8 b =
                    asBf16x8(*reinterpret_cast<const uint4*>(&xs[tdx][nb * 16 + l16][prow + tdy][chunk * 8]));
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i], 0, 0, 0);
            }
        }
    }
    // partial sums of this block: [tap][co][ci] then 32 biases
    float* part = partials + (size_t)blockIdx.x * kP;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int q = wave * 9 + i;
        const int tap = q >> 2, mb = (q >> 1) & 1, nb = q & 1;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int co = mb * 16 + 4 * lg + r, ci = nb * 16 + l16;
            part[(tap * kC + co) * kC + ci] = acc[i][r];
        }
    }
    if (threadIdx.x < 2 * kC) dbPart[threadIdx.x & 1][threadIdx.x >> 1] = dbAcc;
    __syncthreads();
    if (threadIdx.x < kC) part[kW + threadIdx.x] = dbPart[0][threadIdx.x] + dbPart[1][threadIdx.x];
}

__global__ void wgradReduceKernel(const float* __restrict__ partials, int blocks, float* __restrict__ grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kP) return;
    float s = 0.0f;
    for (int b = 0; b < blocks; ++b) s += partials[(size_t)b * kP + i];
    grad[i] = s;
}

// 2x2 mean pooling over the in-range children; one thread per (coarse pixel, 8-channel chunk).
__global__ void poolKernel(const uint16_t* __restrict__ in, int H, int W, uint16_t* __restrict__ out) {
    const int h = (H + 1) / 2, w = (W + 1) / 2;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= h * w * 4) return;
    const int c = i & 3, p = i >> 2;
    const int X = p % w, Y = p / w;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int n = 0;
    for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
            const int yy = 2 * Y + dy, xx = 2 * X + dx;
            if (yy >= H || xx >= W) continue;
            ++n;
            const uint4 v = *reinterpret_cast<const uint4*>(in + ((size_t)yy * W + xx) * kC + c * 8);
            const uint32_t vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                s[2 * j] += bf2f((uint16_t)(vv[j] & 0xffffu));
                s[2 * j + 1] += bf2f((uint16_t)(vv[j] >> 16));
            }
        }
    const float inv = 1.0f / (float)n;
    uint32_t o[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) o[j] = f2bf2(s[2 * j] * inv, s[2 * j + 1] * inv);
    *reinterpret_cast<uint4*>(out + (size_t)p * kC + c * 8) = make_uint4(o[0], o[1], o[2], o[3]);
}

// din[child] (+)= dout[parent] / n(parent); one thread per (fine pixel, chunk).
__global__ void poolBackwardKernel(const uint16_t* __restrict__ dout, int H, int W, uint16_t* __restrict__ din,
                                   int accumulate) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= H * W * 4) return;
    const int c = i & 3, p = i >> 2;
    const int x = p % W, y = p / W;
    const int w = (W + 1) / 2;
    const int X = x / 2, Y = y / 2;
    const int n = (min(2 * Y + 1, H - 1) - 2 * Y + 1) * (min(2 * X + 1, W - 1) - 2 * X + 1);
    const float inv = 1.0f / (float)n;
    const uint4 g = *reinterpret_cast<const uint4*>(dout + ((size_t)Y * w + X) * kC + c * 8);
    uint4 prev = make_uint4(0u, 0u, 0u, 0u);
    uint16_t* dst = din + (size_t)p * kC + c * 8;
    if (accumulate) prev = *reinterpret_cast<const uint4*>(dst);
    const uint32_t gg[4] = {g.x, g.y, g.z, g.w}, pp[4] = {prev.x, prev.y, prev.z, prev.w};
    uint32_t o[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float lo = bf2f((uint16_t)(pp[j] & 0xffffu)) + bf2f((uint16_t)(gg[j] & 0xffffu)) * inv;
        const float hi = bf2f((uint16_t)(pp[j] >> 16)) + bf2f((uint16_t)(gg[j] >> 16)) * inv;
        o[j] = f2bf2(lo, hi);
    }
    *reinterpret_cast<uint4*>(dst) = make_uint4(o[0], o[1], o[2], o[3]);
}

// out = skip + low[y / 2][x / 2]
__global__ void upAddKernel(const uint16_t* __restrict__ skip, const uint16_t* __restrict__ low, int H, int W,
                            uint16_t* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= H * W * 4) return;
    const int c = i & 3, p = i >> 2;
    const int x = p % W, y = p / W;
    const int w = (W + 1) / 2;
    const uint4 a = *reinterpret_cast<const uint4*>(skip + (size_t)p * kC + c * 8);
    const uint4 b = *reinterpret_cast<const uint4*>(low + ((size_t)(y / 2) * w + x / 2) * kC + c * 8);
    const uint32_t aa[4] = {a.x, a.y, a.z, a.w}, bb[4] = {b.x, b.y, b.z, b.w};
    uint32_t o[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float lo = bf2f((uint16_t)(aa[j] & 0xffffu)) + bf2f((uint16_t)(bb[j] & 0xffffu));
        const float hi = bf2f((uint16_t)(aa[j] >> 16)) + bf2f((uint16_t)(bb[j] >> 16));
        o[j] = f2bf2(lo, hi);
    }
    *reinterpret_cast<uint4*>(out + (size_t)p * kC + c * 8) = make_uint4(o[0], o[1], o[2], o[3]);
}

// dlow[Y][X] (+)= sum of the in-range dout children
__global__ void upBackwardKernel(const uint16_t* __restrict__ dout, int H, int W, uint16_t* __restrict__ dlow,
                                 int accumulate) {
    const int h = (H + 1) / 2, w = (W + 1) / 2;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= h * w * 4) return;
    const int c = i & 3, p = i >> 2;
    const int X = p % w, Y = p / w;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
            const int yy = 2 * Y + dy, xx = 2 * X + dx;
            if (yy >= H || xx >= W) continue;
            const uint4 v = *reinterpret_cast<const uint4*>(dout + ((size_t)yy * W + xx) * kC + c * 8);
            const uint32_t vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                s[2 * j] += bf2f((uint16_t)(vv[j] & 0xffffu));
                s[2 * j + 1] += bf2f((uint16_t)(vv[j] >> 16));
            }
        }
    uint16_t* dst = dlow + (size_t)p * kC + c * 8;
    if (accumulate) {
        const uint4 prev = *reinterpret_cast<const uint4*>(dst);
        const uint32_t pp[4] = {prev.x, prev.y, prev.z, prev.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            s[2 * j] = bf2f((uint16_t)(pp[j] & 0xffffu)) + s[2 * j];
            s[2 * j + 1] = bf2f((uint16_t)(pp[j] >> 16)) + s[2 * j + 1];
        }
    }
    uint32_t o[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) o[j] = f2bf2(s[2 * j], s[2 * j + 1]);
    *reinterpret_cast<uint4*>(dst) = make_uint4(o[0], o[1], o[2], o[3]);
}

__global__ void applyKernel(const float* __restrict__ logits, const float* __restrict__ radiance,
                            const float* __restrict__ coarse, int H, int W, float* __restrict__ out,
                            float* __restrict__ err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W * H) return;
    const int x = i % W, y = i / W;
    float lg[28];
    loadLogits(logits, i, lg);
    const f3 o = composePixel(lg, radiance, coarse, W, H, x, y);
    out[3 * i + 0] = o.x;
    out[3 * i + 1] = o.y;
    out[3 * i + 2] = o.z;
    if (err) err[i] = lg[SGI_KPCN_ERR_CHANNEL];
}

// Backward of one reconstruction level at pixel i: the logits' gradient
// (filter taps through the softmax, the blend through the sigmoid, the error
// channel from derr) and the gradient of up(coarse).
__global__ void applyBackwardKernel(const float* __restrict__ logits, const float* __restrict__ radiance,
                                    const float* __restrict__ coarse, const float* __restrict__ dout,
                                    const float* __restrict__ derr, int H, int W, uint16_t* __restrict__ dlogits,
                                    float* __restrict__ dup) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W * H) return;
    const int x = i % W, y = i / W;
    float lg[28], wgt[SGI_KPCN_TAPS];
    loadLogits(logits, i, lg);
    softmax25(lg, wgt);
    const f3 g = ld3(dout + 3 * (size_t)i);
    f3 df = g;
    float dl[32];
#pragma unroll
    for (int j = 0; j < 32; ++j) dl[j] = 0.0f;
    if (coarse) {
        const f3 f = filtered(radiance, W, H, x, y, wgt);
        const f3 u = upBilinear(coarse, (H + 1) / 2, (W + 1) / 2, x, y);
        const float b = sigmoidf(lg[SGI_KPCN_BLEND_CHANNEL]);
        df = g * (1.0f - b);
        const f3 du = g * b;
        dup[3 * (size_t)i + 0] = du.x;
        dup[3 * (size_t)i + 1] = du.y;
        dup[3 * (size_t)i + 2] = du.z;
        dl[SGI_KPCN_BLEND_CHANNEL] = dot(g, u - f) * b * (1.0f - b);
    }
    float dw[SGI_KPCN_TAPS];
    float sdw = 0.0f;
#pragma unroll
    for (int j = 0; j < SGI_KPCN_TAPS; ++j) {
        const int xx = min(max(x + j % 5 - 2, 0), W - 1), yy = min(max(y + j / 5 - 2, 0), H - 1);
        dw[j] = dot(df, ld3(radiance + 3 * ((size_t)yy * W + xx)));
        sdw += wgt[j] * dw[j];
    }
#pragma unroll
    for (int j = 0; j < SGI_KPCN_TAPS; ++j) dl[j] = wgt[j] * (dw[j] - sdw);
    if (derr) dl[SGI_KPCN_ERR_CHANNEL] = derr[i];
    uint4* dst = reinterpret_cast<uint4*>(dlogits + (size_t)i * kC);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        uint32_t o4[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
            o4[j] = f2bf2(dl[q * 8 + 2 * j], dl[q * 8 + 2 * j + 1]);
        dst[q] = make_uint4(o4[0], o4[1], o4[2], o4[3]);
    }
}

// Transpose of up(): coarse pixel (X, Y) collects the fine pixels whose
// bilinear taps land on it (fine rows / columns 2Y - 2 .. 2Y + 3 cover them
// all, clamping included), each with its tap weight.
__global__ void upBilinearBackwardKernel(const float* __restrict__ dup, int H, int W, float* __restrict__ dcoarse) {
    const int h = (H + 1) / 2, w = (W + 1) / 2;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= h * w) return;
    const int X = i % w, Y = i / w;
    f3 acc = splat(0.0f);
    for (int y = max(2 * Y - 2, 0); y <= min(2 * Y + 3, H - 1); ++y) {
        int y0, y1;
        float wy0, wy1;
        bilinearAxis(y, h, y0, y1, wy0, wy1);
        const float wy = (y0 == Y ? wy0 : 0.0f) + (y1 == Y ? wy1 : 0.0f);
        if (wy == 0.0f) continue;
        for (int x = max(2 * X - 2, 0); x <= min(2 * X + 3, W - 1); ++x) {
            int x0, x1;
            float wx0, wx1;
            bilinearAxis(x, w, x0, x1, wx0, wx1);
            const float wx = (x0 == X ? wx0 : 0.0f) + (x1 == X ? wx1 : 0.0f);
            if (wx == 0.0f) continue;
            acc = acc + ld3(dup + 3 * ((size_t)y * W + x)) * (wy * wx);
        }
    }
    dcoarse[3 * (size_t)i + 0] = acc.x;
    dcoarse[3 * (size_t)i + 1] = acc.y;
    dcoarse[3 * (size_t)i + 2] = acc.z;
}

// The radiance pyramid, one 16x16 level-0 tile per block: the tile's 8x8,
// 4x4, 2x2 and 1x1 pixels of levels 1..4 (a level-l pixel's children all lie
// in the same tile), each the mean of its in-range children in the order
// (0,0), (0,1), (1,0), (1,1) as poolKernel sums.
// Levels 1..4 of one 16x16 tile's pyramid from its level-0 radiance in buf[0].
__device__ __forceinline__ void tilePyramid(float (*buf)[kTS * kTS * 3], int tx, int ty, int H, int W,
                                            float* __restrict__ pyr) {
    int hp = H, wp = W;  // the finer level's size
    size_t off = 0;      // this level's offset in pyr (floats)
    int side = kTS;      // the finer level's tile edge
    for (int l = 1; l < SGI_KPCN_LEVELS; ++l) {
        __syncthreads();
        const int h = (hp + 1) / 2, w = (wp + 1) / 2, s = side / 2;
        const float* src = buf[(l - 1) & 1];
        float* dst = buf[l & 1];
        if ((int)threadIdx.x < s * s) {
            const int lx = (int)threadIdx.x % s, ly = (int)threadIdx.x / s;
            const int X = tx * s + lx, Y = ty * s + ly;
            if (X < w && Y < h) {
                float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
                int n = 0;
                for (int dy = 0; dy < 2; ++dy)
                    for (int dx = 0; dx < 2; ++dx) {
                        if (2 * Y + dy >= hp || 2 * X + dx >= wp) continue;
                        const int c = (2 * ly + dy) * side + 2 * lx + dx;
                        a0 += src[3 * c + 0];
                        a1 += src[3 * c + 1];
                        a2 += src[3 * c + 2];
                        ++n;
                    }
                const float inv = n == 4 ? 0.25f : (n == 2 ? 0.5f : 1.0f);  // 1 / n, exact
                const int d = ly * s + lx;
                dst[3 * d + 0] = a0 * inv;
                dst[3 * d + 1] = a1 * inv;
                dst[3 * d + 2] = a2 * inv;
                float* o = pyr + off + 3 * ((size_t)Y * w + X);
                o[0] = a0 * inv;
                o[1] = a1 * inv;
                o[2] = a2 * inv;
            }
        }
        off += (size_t)h * w * 3;
        hp = h;
        wp = w;
        side = s;
    }
}

__global__ __launch_bounds__(256) void pyramidKernel(const float* __restrict__ r0, int H, int W, int tilesX,
                                                     float* __restrict__ pyr) {
    __shared__ float buf[2][kTS * kTS * 3];
    const int tx = blockIdx.x % tilesX, ty = blockIdx.x / tilesX;
    const int x = tx * kTS + (int)threadIdx.x % kTS, y = ty * kTS + (int)threadIdx.x / kTS;
    if (x < W && y < H) {
        const f3 v = ld3(r0 + 3 * ((size_t)y * W + x));
        buf[0][3 * threadIdx.x + 0] = v.x;
        buf[0][3 * threadIdx.x + 1] = v.y;
        buf[0][3 * threadIdx.x + 2] = v.z;
    }
    tilePyramid(buf, tx, ty, H, W, pyr);
}

// The inference features ([h][w][16]) and the radiance pyramid in one pass: one
// 16x16 tile per block, the tile's radiance kept in LDS for its pyramid.
__global__ __launch_bounds__(256) void featuresPyramidKernel(GbufArgs g, FrameArgs fr,
                                                             const float4* __restrict__ reservoir, int tilesX,
                                                             uint16_t* __restrict__ feat16,
                                                             float* __restrict__ radiance, float* __restrict__ pyr) {
    __shared__ float buf[2][kTS * kTS * 3];
    const int tx = blockIdx.x % tilesX, ty = blockIdx.x / tilesX;
    const int x = tx * kTS + (int)threadIdx.x % kTS, y = ty * kTS + (int)threadIdx.x / kTS;
    if (x < g.W && y < g.H) {
        uint32_t o[16];
        float L[3];
        pixelFeatures(g, fr, reservoir, x, y, o, L);
        const size_t i = (size_t)y * g.W + x;
        radiance[3 * i + 0] = L[0];
        radiance[3 * i + 1] = L[1];
        radiance[3 * i + 2] = L[2];
        uint4* dst = reinterpret_cast<uint4*>(feat16 + i * 16);
        dst[0] = make_uint4(o[0], o[1], o[2], o[3]);
        dst[1] = make_uint4(o[4], o[5], o[6], o[7]);
        buf[0][3 * threadIdx.x + 0] = L[0];
        buf[0][3 * threadIdx.x + 1] = L[1];
        buf[0][3 * threadIdx.x + 2] = L[2];
    }
    tilePyramid(buf, tx, ty, g.H, g.W, pyr);
}

// Loss on the reconstruction: the relative squared error of out and the error
// head's regression onto log(r_p + 1e-4); gradients of out and of the error
// logit, times scale.
// alt: NULL (the error channel learns log(r_p + 1e-4)) or another output of the
// same input (the error channel learns log(r_p + 1e-4) - log(a_p + 1e-4), a_p
// its relative squared error, so exp(err) estimates the ratio of the two).
__global__ __launch_bounds__(256) void lossKernel(const float* __restrict__ out, const float* __restrict__ logits,
                                                  const float* __restrict__ target, const float* __restrict__ alt,
                                                  int H, int W, float errWeight, float scale,
                                                  float* __restrict__ dout, float* __restrict__ derr,
                                                  float* __restrict__ lossSums) {
    __shared__ float red[2][256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l0 = 0.0f, l1 = 0.0f;
    if (i < W * H) {
        const float invN = 1.0f / (3.0f * (float)W * (float)H), invP = 1.0f / ((float)W * (float)H);
        const f3 o = ld3(out + 3 * (size_t)i);
        const f3 t = ld3(target + 3 * (size_t)i);
        const f3 den = f3{t.x * t.x + 0.01f, t.y * t.y + 0.01f, t.z * t.z + 0.01f};
        const f3 e = o - t;
        const float rx = e.x * e.x / den.x, ry = e.y * e.y / den.y, rz = e.z * e.z / den.z;
        l0 = rx + ry + rz;
        const float rp = l0 * (1.0f / 3.0f);
        float errT = logf(rp + 1e-4f);
        if (alt) {
            const f3 ea = ld3(alt + 3 * (size_t)i) - t;
            const float ra = (ea.x * ea.x / den.x + ea.y * ea.y / den.y + ea.z * ea.z / den.z) * (1.0f / 3.0f);
            errT -= logf(ra + 1e-4f);
        }
        const float de = logits[(size_t)i * kC + SGI_KPCN_ERR_CHANNEL] - errT;
        l1 = de * de;
        const f3 d = f3{2.0f * e.x / den.x, 2.0f * e.y / den.y, 2.0f * e.z / den.z} * (invN * scale);
        dout[3 * (size_t)i + 0] = d.x;
        dout[3 * (size_t)i + 1] = d.y;
        dout[3 * (size_t)i + 2] = d.z;
        derr[i] = 2.0f * errWeight * de * invP * scale;
    }
    red[0][threadIdx.x] = l0;
    red[1][threadIdx.x] = l1;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(&lossSums[0], red[0][0]);
        atomicAdd(&lossSums[1], red[1][0]);
    }
}

__global__ void adamKernel(float* __restrict__ params, const float* __restrict__ grad, float* __restrict__ m,
                           float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= SGI_KPCN_PARAMS) return;
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    params[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// BF16 forward weights, transposed / mirrored input-gradient weights and FP32 biases.
__global__ void weightCopiesKernel(const float* __restrict__ params, uint16_t* __restrict__ w,
                                   uint16_t* __restrict__ wt, float* __restrict__ biases) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= SGI_KPCN_LAYERS * kP) return;
    const int l = i / kP, r = i % kP;
    const float val = params[i];
    if (r >= kW) {
        if (biases) biases[l * kC + (r - kW)] = val;
        return;
    }
    const int t = r / (kC * kC), o = (r / kC) % kC, in = r % kC;
    if (w) w[(size_t)l * kW + r] = f2bf(val);
    if (wt) wt[(size_t)l * kW + ((8 - t) * kC + in) * kC + o] = f2bf(val);
}

// Per tile (one thread each): whether a flagged tile lies within two tiles,
// and if so the tile's index appended to dilatedList (one atomic per wave;
// the list's order does not matter, every listed tile is processed on its
// own); counts[0] (may be NULL) gathers the flagged tiles.  Both counters
// start at 0 (the level-0 head's tile 0 clears them).
__global__ __launch_bounds__(256) void dilateListKernel(const uint8_t* __restrict__ flags, int tilesX, int tilesY,
                                                        int32_t* __restrict__ dilatedList,
                                                        int32_t* __restrict__ counts) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = tilesX * tilesY;
    bool f = false, d = false;
    if (t < n) {
        f = flags[t] != 0;
        const int tx = t % tilesX, ty = t / tilesX;
        for (int dy = -2; dy <= 2; ++dy)
            for (int dx = -2; dx <= 2; ++dx) {
                const int xx = tx + dx, yy = ty + dy;
                if (xx >= 0 && xx < tilesX && yy >= 0 && yy < tilesY && flags[yy * tilesX + xx]) d = true;
            }
    }
    const int lane = (int)threadIdx.x & 63;
    const uint64_t bd = __ballot(d), bf = __ballot(f);
    int base = 0;
    if (lane == 0 && bd) base = atomicAdd(&dilatedList[0], __popcll(bd));
    base = __shfl(base, 0);
    if (d) dilatedList[1 + base + __popcll(bd & ((1ull << lane) - 1ull))] = t;
    if (counts && lane == 0 && bf) atomicAdd(&counts[0], __popcll(bf));
}

__global__ void clearCountsKernel(int n, int32_t* __restrict__ counts) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) counts[i] = 0;
}

int gridFor(int n, int b) { return (n + b - 1) / b; }

}  // namespace

namespace sgi_kpcn {

template <int IN, int OUT, int AIN, int AOUT>
hipError_t launchConv(const ConvSrc& src, const void* w, const float* bias, int relu, int H, int W, void* out,
                      float* err, hipStream_t s) {
    const int tilesX = (W + kTS - 1) / kTS, nTiles = tilesX * ((H + kTS - 1) / kTS);
    // convBlocksPerCu blocks per CU stay resident (LDS and VGPRs), so 256 times that many cover
    // the 256 CUs once and each walks a contiguous run of tiles with its next halo in flight
    const int resident = 256 * convBlocksPerCu<IN, OUT, AIN>();
    const int grid = nTiles < resident ? nTiles : resident;
    hipLaunchKernelGGL((conv3x3Kernel<IN, OUT, AIN, AOUT>), dim3(grid), dim3(256), 0, s, src, w, bias, relu, H, W,
                       tilesX, nTiles, out, err);
    return hipGetLastError();
}

// The same form with FP8 (f8 != 0) or BF16 activations in and out.
template <int IN, int OUT>
hipError_t launchConvAct(int f8, const ConvSrc& src, const void* w, const float* bias, int relu, int H, int W,
                         void* out, float* err, hipStream_t s) {
    if (f8) return launchConv<IN, OUT, ACT_F8, ACT_F8>(src, w, bias, relu, H, W, out, err, s);
    return launchConv<IN, OUT, ACT_BF16, ACT_BF16>(src, w, bias, relu, H, W, out, err, s);
}

ConvSrc convSrc(const void* a, const void* b = nullptr) {
    ConvSrc src{};
    src.a = static_cast<const uint16_t*>(a);
    src.b = static_cast<const uint16_t*>(b);
    return src;
}

hipError_t conv3x3(const uint16_t* in, const uint16_t* mask, const uint16_t* w, const float* bias, int relu, int H,
                   int W, void* out, int outF32, hipStream_t s) {
    ConvSrc src = convSrc(in);
    src.mask = mask;
    if (outF32) return launchConv<IN_PLAIN, OUT_F32, ACT_BF16, ACT_BF16>(src, w, bias, relu, H, W, out, nullptr, s);
    return launchConv<IN_PLAIN, OUT_BF16, ACT_BF16, ACT_BF16>(src, w, bias, relu, H, W, out, nullptr, s);
}

hipError_t convF8(const void* in, int H, int W, const void* w, const float* scales, void* out, hipStream_t s) {
    return launchConv<IN_PLAIN, OUT_BF16, ACT_F8, ACT_F8>(convSrc(in), w, scales, 1, H, W, out, nullptr, s);
}

hipError_t conv3x3In16AndPool(const uint16_t* in16, int H, int W, const uint16_t* w, const float* bias, void* out,
                              void* pooled, int f8out, hipStream_t s) {
    ConvSrc src = convSrc(in16);
    src.pooled = static_cast<uint16_t*>(pooled);
    if (f8out)
        return launchConv<IN_PLAIN16, OUT_BF16_POOL, ACT_BF16, ACT_F8>(src, w, bias, 1, H, W, out, nullptr, s);
    return launchConv<IN_PLAIN16, OUT_BF16_POOL, ACT_BF16, ACT_BF16>(src, w, bias, 1, H, W, out, nullptr, s);
}

hipError_t convUpAddApply(const void* skip, const void* low, int H, int W, const void* w, const float* bias,
                          const float* radiance, const float* coarse, float* out, float* err, int f8,
                          hipStream_t s, float threshold, uint8_t* flags, int32_t* dilatedList, int32_t* counts) {
    ConvSrc src = convSrc(skip, low);
    src.radiance = radiance;
    src.coarse = coarse;
    src.flags = flags;
    src.dilatedList = dilatedList;
    src.flagCounts = counts;
    src.flagThreshold = threshold;
    return launchConvAct<IN_UPADD, OUT_APPLY>(f8, src, w, bias, 0, H, W, out, err, s);
}

hipError_t convAndPool(const void* in, int H, int W, const void* w, const float* bias, void* out, void* pooled,
                       int f8, hipStream_t s) {
    ConvSrc src = convSrc(in);
    src.pooled = static_cast<uint16_t*>(pooled);
    return launchConvAct<IN_PLAIN, OUT_BF16_POOL>(f8, src, w, bias, 1, H, W, out, nullptr, s);
}

hipError_t convUpAdd(const void* skip, const void* low, int H, int W, const void* w, const float* bias, void* out,
                     int f8, hipStream_t s) {
    return launchConvAct<IN_UPADD, OUT_BF16>(f8, convSrc(skip, low), w, bias, 1, H, W, out, nullptr, s);
}

hipError_t convApply(const void* in, int H, int W, const void* w, const float* bias, const float* radiance,
                     const float* coarse, float* out, float* err, int f8, hipStream_t s) {
    ConvSrc src = convSrc(in);
    src.radiance = radiance;
    src.coarse = coarse;
    return launchConvAct<IN_PLAIN, OUT_APPLY>(f8, src, w, bias, 0, H, W, out, err, s);
}

hipError_t encodeBf16(const float* in, int64_t n, uint16_t* out, hipStream_t s) {
    hipLaunchKernelGGL(encodeBf16Kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, n, out);
    return hipGetLastError();
}

hipError_t encodeF8(const float* in, int64_t n, uint8_t* out, hipStream_t s) {
    const int64_t words = (n + 3) / 4;
    hipLaunchKernelGGL(encodeF8Kernel, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s, in, n, out);
    return hipGetLastError();
}

int wgradBlocks(int H, int W) {
    const int nTiles = ((W + kTS - 1) / kTS) * ((H + kTS - 1) / kTS);
    return nTiles < 512 ? nTiles : 512;
}

hipError_t wgrad(const uint16_t* x, const uint16_t* dy, const uint16_t* mask, int H, int W, float* partials,
                 float* grad, hipStream_t s) {
    const int tilesX = (W + kTS - 1) / kTS, nTiles = tilesX * ((H + kTS - 1) / kTS);
    const int blocks = wgradBlocks(H, W);
    hipLaunchKernelGGL(wgradKernel, dim3(blocks), dim3(256), 0, s, x, dy, mask, H, W, tilesX, nTiles, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(wgradReduceKernel, dim3(gridFor(kP, 256)), dim3(256), 0, s, partials, blocks, grad);
    return hipGetLastError();
}

hipError_t pool(const uint16_t* in, int H, int W, uint16_t* out, hipStream_t s) {
    const int n = ((H + 1) / 2) * ((W + 1) / 2) * 4;
    hipLaunchKernelGGL(poolKernel, dim3(gridFor(n, 256)), dim3(256), 0, s, in, H, W, out);
    return hipGetLastError();
}

hipError_t poolBackward(const uint16_t* dout, int H, int W, uint16_t* din, int accumulate, hipStream_t s) {
    const int n = H * W * 4;
    hipLaunchKernelGGL(poolBackwardKernel, dim3(gridFor(n, 256)), dim3(256), 0, s, dout, H, W, din, accumulate);
    return hipGetLastError();
}

hipError_t upAdd(const uint16_t* skip, const uint16_t* low, int H, int W, uint16_t* out, hipStream_t s) {
    const int n = H * W * 4;
    hipLaunchKernelGGL(upAddKernel, dim3(gridFor(n, 256)), dim3(256), 0, s, skip, low, H, W, out);
    return hipGetLastError();
}

hipError_t upBackward(const uint16_t* dout, int H, int W, uint16_t* dlow, int accumulate, hipStream_t s) {
    const int n = ((H + 1) / 2) * ((W + 1) / 2) * 4;
    hipLaunchKernelGGL(upBackwardKernel, dim3(gridFor(n, 256)), dim3(256), 0, s, dout, H, W, dlow, accumulate);
    return hipGetLastError();
}

size_t pyramidFloats(int H, int W) {
    size_t n = 0;
    for (int l = 1; l < SGI_KPCN_LEVELS; ++l) {
        H = (H + 1) / 2;
        W = (W + 1) / 2;
        n += (size_t)H * W * 3;
    }
    return n;
}

hipError_t pyramid(const float* radiance, int H, int W, float* pyr, hipStream_t s) {
    const int tilesX = (W + kTS - 1) / kTS, nTiles = tilesX * ((H + kTS - 1) / kTS);
    hipLaunchKernelGGL(pyramidKernel, dim3(nTiles), dim3(256), 0, s, radiance, H, W, tilesX, pyr);
    return hipGetLastError();
}

hipError_t applyBackward(const float* logits, const float* radiance, const float* coarse, const float* dout,
                         const float* derr, int H, int W, uint16_t* dlogits, float* dup, hipStream_t s) {
    hipLaunchKernelGGL(applyBackwardKernel, dim3(gridFor(W * H, 256)), dim3(256), 0, s, logits, radiance, coarse,
                       dout, derr, H, W, dlogits, dup);
    return hipGetLastError();
}

hipError_t upBilinearBackward(const float* dup, int H, int W, float* dcoarse, hipStream_t s) {
    const int n = ((H + 1) / 2) * ((W + 1) / 2);
    hipLaunchKernelGGL(upBilinearBackwardKernel, dim3(gridFor(n, 256)), dim3(256), 0, s, dup, H, W, dcoarse);
    return hipGetLastError();
}

hipError_t features(const GbufArgs& g, const FrameArgs& fr, const float* reservoir, uint16_t* feat, float* radiance,
                    hipStream_t s) {
    hipLaunchKernelGGL(featuresKernel<32>, dim3(gridFor(g.W * g.H, 256)), dim3(256), 0, s, g, fr,
                       (const float4*)reservoir, feat, radiance);
    return hipGetLastError();
}

hipError_t features16Pyramid(const GbufArgs& g, const FrameArgs& fr, const float* reservoir, uint16_t* feat16,
                             float* radiance, float* pyr, hipStream_t s) {
    const int tilesX = (g.W + kTS - 1) / kTS, nTiles = tilesX * ((g.H + kTS - 1) / kTS);
    hipLaunchKernelGGL(featuresPyramidKernel, dim3(nTiles), dim3(256), 0, s, g, fr, (const float4*)reservoir, tilesX,
                       feat16, radiance, pyr);
    return hipGetLastError();
}

hipError_t apply(const float* logits, const float* radiance, const float* coarse, int H, int W, float* out,
                 float* err, hipStream_t s) {
    hipLaunchKernelGGL(applyKernel, dim3(gridFor(W * H, 256)), dim3(256), 0, s, logits, radiance, coarse, H, W, out,
                       err);
    return hipGetLastError();
}

hipError_t loss(const float* out, const float* logits, const float* target, const float* alt, int H, int W,
                float errWeight, float scale, float* dout, float* derr, float* lossSums, hipStream_t s) {
    hipLaunchKernelGGL(lossKernel, dim3(gridFor(W * H, 256)), dim3(256), 0, s, out, logits, target, alt, H, W, errWeight,
                       scale, dout, derr, lossSums);
    return hipGetLastError();
}

hipError_t adam(float* params, const float* grad, float* m, float* v, float lr, float b1, float b2, float eps,
                int step, uint16_t* w, uint16_t* wt, float* biases, hipStream_t s) {
    const float c1 = 1.0f / (1.0f - powf(b1, (float)step)), c2 = 1.0f / (1.0f - powf(b2, (float)step));
    hipLaunchKernelGGL(adamKernel, dim3(gridFor(SGI_KPCN_PARAMS, 256)), dim3(256), 0, s, params, grad, m, v, lr, b1,
                       b2, eps, c1, c2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(weightCopiesKernel, dim3(gridFor(SGI_KPCN_PARAMS, 256)), dim3(256), 0, s, params, w, wt,
                       biases);
    return hipGetLastError();
}

hipError_t packWeights(const float* params, uint16_t* w, uint16_t* wt, float* biases, hipStream_t s) {
    hipLaunchKernelGGL(weightCopiesKernel, dim3(gridFor(SGI_KPCN_PARAMS, 256)), dim3(256), 0, s, params, w, wt,
                       biases);
    return hipGetLastError();
}

hipError_t dilateFlags(int H, int W, const uint8_t* flags, int32_t* dilatedList, int32_t* counts, hipStream_t s) {
    const int tilesX = (W + 15) / 16, tilesY = (H + 15) / 16, n = tilesX * tilesY;
    hipLaunchKernelGGL(dilateListKernel, dim3(gridFor(n, 256)), dim3(256), 0, s, flags, tilesX, tilesY, dilatedList,
                       counts);
    return hipGetLastError();
}

hipError_t clearCounts(int H, int W, int32_t* counts, hipStream_t s) {
    const int n = ((W + 15) / 16) * ((H + 15) / 16);
    hipLaunchKernelGGL(clearCountsKernel, dim3(gridFor(n + 1, 256)), dim3(256), 0, s, n, counts);
    return hipGetLastError();
}

}  // namespace sgi_kpcn
