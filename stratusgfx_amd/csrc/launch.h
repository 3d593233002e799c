// Host launchers of the GI kernels (each defined next to its kernel).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/sgi_kpcn.h"
#include "gi_device.h"
#include "views.h"

namespace sgi_launch {

// flags: device scratch of n bytes
hipError_t cullGlobal(const sgi_dev::CsmArgs& csm, const sgi_dev::FrameArgs& fr, const sgi_vpl* in, int n,
                      sgi_vpl* out, int32_t* visible, uint8_t* flags, hipStream_t s);

hipError_t gatherVisibleEntries(const sgi_atlas_entry* all, const int32_t* visible, int maxCount,
                                sgi_atlas_entry* out, hipStream_t s);

// taa.fs: current frame + history -> anti-aliased frame (velocity may be NULL)
hipError_t taa(const float* screen, const float* prev, const float* velocity, int W, int H, float* out, hipStream_t s);

// lit [n][3] FP32 -> ACES film + gamma -> RGBA8 texels
hipError_t tonemap(const float* lit, long n, float gamma, uint32_t* out, hipStream_t s);

// VPL cube producer (shadowVpl.fs through the reference's cube update loop): the
// cubes of vpls[updates[0 .. numUpdates)] rendered from `count` world-space
// triangles, in batches of batchMax lights; scratch holds cubeListInts(batchMax, count) ints.
size_t cubeListInts(int batch, int count);
hipError_t renderVplCubes(const float* pos, const uint32_t* colors, int count, int cull, const sgi_vpl* vpls,
                          const int32_t* updates, int numUpdates, const sgi_atlas_entry* entries, float nearPlane,
                          const sgi_dev::CubeArgs& dist, int format, const sgi_dev::CubeArgs& diffuse,
                          bool withDiffuse, int batchMax, int32_t* scratch, hipStream_t s);

// Sun cascades' depth render (RenderCSMDepth_ with csm.vs): layers [4][res][res];
// scratch holds csmScratchFloats(count) floats.  keepSign: +1 keep NDC-counter-clockwise
// triangles, -1 clockwise, 0 both.
size_t csmScratchFloats(int count);
hipError_t renderCsm(const float* pos, int count, int keepSign, float slope, int res,
                     const float (*projViews)[16], float* scratch, float* depth, hipStream_t s);

// PerformTaaPostFx_'s history copy: TAA output -> the RGB8 history as read back (k / 255)
hipError_t taaHistory(const float* taa, long n, float* prev, hipStream_t s);

// fxaa_luminance.fs + fxaa_smoothing.fs fused: tonemapped RGBA8 -> anti-aliased RGBA8
hipError_t fxaa(const uint32_t* in, int W, int H, float contrastThreshold, float relativeThreshold,
                float subpixelBlending, uint32_t* out, hipStream_t s);

// packed G-buffer attachments -> FP32 planes (velocity may be NULL)
hipError_t gbufferPack(const sgi_gbuffer& planes, const sgi_gbuffer_packed& out, hipStream_t s);
hipError_t gbufferUnpack(const sgi_gbuffer_packed& g, float* normal, float* albedo, float* refl, float* rma,
                         float* velocity, hipStream_t s);

// vplUpdatedData[s] colour and specular position from the diffuse / distance cubes
hipError_t colorVpls(const sgi_dev::CubeArgs& diffuse, const sgi_dev::CubeArgs& dist, int distFormat,
                     const sgi_dev::FrameArgs& fr, float mult, const int32_t* visible, int maxCount,
                     const sgi_atlas_entry* entries, sgi_vpl* vpls, hipStream_t s);

hipError_t cullTiles(const sgi_dev::GbufArgs& g, const sgi_dev::FrameArgs& fr, int tileSize, int tilesX, int tilesY,
                     int K, const sgi_vpl* vpls, const int32_t* visible, int remap, sgi_tile_stage1* stage1,
                     int32_t* tiles, hipStream_t s);

// prep: device scratch of SGI_MAX_TOTAL_VPLS_PER_FRAME * packedVplBytes() (the
// sampled estimators' per-frame VPL records, then their binary16 positions)
size_t packedVplBytes();

hipError_t gather(const sgi_dev::GbufArgs& g, const sgi_dev::FrameArgs& fr, const sgi_dev::CubeArgs& cubes,
                  int format, int tileSize, int tilesX, int tilesY, int K, const sgi_vpl* vpls,
                  const int32_t* visible, const sgi_atlas_entry* entries, const int32_t* tiles, int precision,
                  void* prep, float* color, float* reservoir, hipStream_t s);

// modes / samples / fp64 as sgi_gather_opts
hipError_t gatherEx(const sgi_dev::GbufArgs& g, const sgi_dev::FrameArgs& fr, const sgi_dev::CubeArgs& cubes,
                    int format, int mode, int samples, int rejectionLimit, int accumulate64, int half, const sgi_vpl* vpls,
                    const int32_t* visible, const sgi_atlas_entry* entries, void* prep, float* color,
                    float* reservoir, hipStream_t s);

hipError_t directLight(const sgi_dev::GbufArgs& g, const sgi_dev::FrameArgs& fr, const sgi_dev::CsmArgs& csm,
                       float* lit, hipStream_t s);

struct DenoiseBuffers {
    float4* resA;     // [P] ping
    float4* resB;     // [P] pong
    float4* normals;  // [P] decoded current normal (xyz) and depth (w)
};

hipError_t denoise(const sgi_dev::GbufArgs& cur, const sgi_dev::GbufArgs& prev, const sgi_dev::FrameArgs& fr,
                   const int32_t* visible, const float* screen, const float* color, const float* reservoir,
                   const float* prevGi, const float* prevHistory, float* combined, float* giOut,
                   float* reservoirOut, float* historyOut, const DenoiseBuffers& ws, hipStream_t s);

// Batched views (views: device array of ViewDesc, numbered in item order).
hipError_t cullGlobalViews(const sgi_batch::ViewDesc* views, int numViews, const sgi_vpl* in, int n, hipStream_t s);
hipError_t gatherVisibleEntriesViews(const sgi_batch::ViewDesc* views, int numViews, const sgi_atlas_entry* all,
                                     int maxCount, hipStream_t s);
hipError_t cullTilesViews(const sgi_batch::ViewDesc* views, int numViews, int totalTiles, int tileSize, int K,
                          hipStream_t s);
hipError_t gatherViews(const sgi_batch::ViewDesc* views, int numViews, int totalTiles, int tileSize, int K,
                       const sgi_dev::CubeArgs& cubes, int format, hipStream_t s);
hipError_t directLightViews(const sgi_batch::ViewDesc* views, int numViews, int totalTiles16, hipStream_t s);
hipError_t denoiseViews(const sgi_batch::ViewDesc* views, int numViews, int totalBlocks, hipStream_t s);

// The merge and the three filter passes into reservoirOut; tiles (a device
// list of 16x16 tiles: int32 count, then tile indices; NULL = the whole
// frame) limits them, and the merge's decoded normals, to the listed tiles;
// mergeOnly stops after the merge (whose decoded normals the temporal pass
// reads).
hipError_t denoiseSpatial(const sgi_dev::GbufArgs& cur, const sgi_dev::FrameArgs& fr, const int32_t* visible,
                          const float* reservoir, float* reservoirOut, const DenoiseBuffers& ws, const int32_t* tiles,
                          bool mergeOnly, hipStream_t s);

// The temporal blend and combine on a filtered reservoir (normals from denoiseSpatial).
hipError_t denoiseTemporal(const sgi_dev::GbufArgs& cur, const sgi_dev::GbufArgs& prev, const sgi_dev::FrameArgs& fr,
                           const float* screen, const float* color, const float* filtered, const float* prevGi,
                           const float* prevHistory, float* combined, float* giOut, float* historyOut,
                           const DenoiseBuffers& ws, hipStream_t s);
// The temporal pass on the learned denoiser's output (net, FP32 [h][w][3]):
// also writes reservoirOut = (net, 1); decodes the centre normals itself (no
// merge pass).  fallback (per 16x16 tile, NULL = none): the marked tiles take
// the a-trous result already in reservoirOut instead, with the decoded
// normals of denoiseSpatial (the temporal pass of sgi_denoise).
hipError_t denoiseTemporalNet(const sgi_dev::GbufArgs& cur, const sgi_dev::GbufArgs& prev, const sgi_dev::FrameArgs& fr,
                              const float* screen, const float* color, const float* net, const float* prevGi,
                              const float* prevHistory, float* combined, float* giOut, float* reservoirOut,
                              float* historyOut, const uint8_t* fallback, const float4* normals, hipStream_t s);

// sgi_denoise_learned's work (api.hip) on checked arguments: workspace of
// sgi_kpcn_workspace_bytes(cur.H, cur.W) bytes, dws the a-trous planes of the
// fallback.  Shared with the batched views' learned form.
hipError_t denoiseLearned(const sgi_dev::GbufArgs& cur, const sgi_dev::GbufArgs& prev, const sgi_dev::FrameArgs& fr,
                          const int32_t* visibleIndices, const float* screen, const float* color,
                          const float* reservoir, const float* prevGi, const float* prevHistory,
                          const sgi_kpcn_net& net, float threshold, void* workspace, float* combined, float* giOut,
                          float* reservoirOut, float* historyOut, int32_t* fallbackTiles, const DenoiseBuffers& dws,
                          hipEvent_t screenReady, hipStream_t s);

}  // namespace sgi_launch

// Learned-denoiser launchers (kpcn.hip).
namespace sgi_kpcn {

hipError_t conv3x3(const uint16_t* in, const uint16_t* mask, const uint16_t* w, const float* bias, int relu, int H,
                   int W, void* out, int outF32, hipStream_t s);
// Inference forms of the convolution with the neighbouring pass fused into the
// input staging (upsample + skip) or the epilogue (2x2 pooling of the output;
// one reconstruction level: the 5x5 filter and the blend with the coarser
// output); same results as the separate kernels.
// The first convolution on the inference features' [h][w][16] layout (BF16),
// with the 2x2 mean of its output for the next level; f8out: both outputs FP8
// E4M3 (bias then holds the epilogue's [k1[32], k2[32]]).
hipError_t conv3x3In16AndPool(const uint16_t* in16, int H, int W, const uint16_t* w, const float* bias, void* out,
                              void* pooled, int f8out, hipStream_t s);
// The remaining forms take f8: 0 for BF16 activations and weights with a bias,
// 1 for FP8 E4M3 activations and weights with the epilogue [k1[32], k2[32]].
// The level-0 head: its input e0 + up(d1) staged as convUpAdd stages it, its
// epilogue the reconstruction step as convApply's. With flags set (err too) the
// epilogue also writes the per-tile fallback flags (flag = mean exp(min(err, 80))
// over the tile > threshold), the counts record and the list counter's reset.
hipError_t convUpAddApply(const void* skip, const void* low, int H, int W, const void* w, const float* bias,
                          const float* radiance, const float* coarse, float* out, float* err, int f8,
                          hipStream_t s, float threshold = 0.0f, uint8_t* flags = nullptr,
                          int32_t* dilatedList = nullptr, int32_t* counts = nullptr);
hipError_t convAndPool(const void* in, int H, int W, const void* w, const float* bias, void* out, void* pooled,
                       int f8, hipStream_t s);
hipError_t convUpAdd(const void* skip, const void* low, int H, int W, const void* w, const float* bias, void* out,
                     int f8, hipStream_t s);
hipError_t convApply(const void* in, int H, int W, const void* w, const float* bias, const float* radiance,
                     const float* coarse, float* out, float* err, int f8, hipStream_t s);
// FP8 in, FP8 out (ReLU): out = e4m3(max(fma(acc, k1, k2), 0)), scales = [k1[32], k2[32]].
hipError_t convF8(const void* in, int H, int W, const void* w, const float* scales, void* out, hipStream_t s);
hipError_t encodeBf16(const float* in, int64_t n, uint16_t* out, hipStream_t s);
hipError_t encodeF8(const float* in, int64_t n, uint8_t* out, hipStream_t s);
int wgradBlocks(int H, int W);
hipError_t wgrad(const uint16_t* x, const uint16_t* dy, const uint16_t* mask, int H, int W, float* partials,
                 float* grad, hipStream_t s);
hipError_t pool(const uint16_t* in, int H, int W, uint16_t* out, hipStream_t s);
hipError_t poolBackward(const uint16_t* dout, int H, int W, uint16_t* din, int accumulate, hipStream_t s);
hipError_t upAdd(const uint16_t* skip, const uint16_t* low, int H, int W, uint16_t* out, hipStream_t s);
hipError_t upBackward(const uint16_t* dout, int H, int W, uint16_t* dlow, int accumulate, hipStream_t s);
size_t pyramidFloats(int H, int W);
hipError_t pyramid(const float* radiance, int H, int W, float* pyr, hipStream_t s);
hipError_t applyBackward(const float* logits, const float* radiance, const float* coarse, const float* dout,
                         const float* derr, int H, int W, uint16_t* dlogits, float* dup, hipStream_t s);
hipError_t upBilinearBackward(const float* dup, int H, int W, float* dcoarse, hipStream_t s);
hipError_t features(const sgi_dev::GbufArgs& g, const sgi_dev::FrameArgs& fr, const float* reservoir, uint16_t* feat,
                    float* radiance, hipStream_t s);
// The same features in the inference layout [h][w][16], and the radiance
// pyramid (levels 1..4, the layout of pyramid()) from the same pass.
hipError_t features16Pyramid(const sgi_dev::GbufArgs& g, const sgi_dev::FrameArgs& fr, const float* reservoir,
                             uint16_t* feat16, float* radiance, float* pyr, hipStream_t s);
hipError_t apply(const float* logits, const float* radiance, const float* coarse, int H, int W, float* out,
                 float* err, hipStream_t s);
hipError_t loss(const float* out, const float* logits, const float* target, const float* alt, int H, int W,
                float errWeight, float scale, float* dout, float* derr, float* lossSums, hipStream_t s);
hipError_t adam(float* params, const float* grad, float* m, float* v, float lr, float b1, float b2, float eps,
                int step, uint16_t* w, uint16_t* wt, float* biases, hipStream_t s);
hipError_t packWeights(const float* params, uint16_t* w, uint16_t* wt, float* biases, hipStream_t s);
// From the per-tile flags the level-0 head formed (convUpAddApply with flags): the
// list of the flagged tiles grown by two tiles (the a-trous passes' 28-pixel reach;
// int32 [1 + tiles]: count, then the tile indices in no particular order) and, when
// counts is set, the flagged-tile count completing the sgi_denoise_learned
// fallbackTiles record (count, then one flag per tile).
hipError_t dilateFlags(int H, int W, const uint8_t* flags, int32_t* dilatedList, int32_t* counts, hipStream_t s);
// counts[0] = 0 and counts[1 + t] = 0 for every tile (no tile falls back).
hipError_t clearCounts(int H, int W, int32_t* counts, hipStream_t s);

}  // namespace sgi_kpcn
