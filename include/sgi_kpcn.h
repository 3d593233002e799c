/*
 * sgi_kpcn.h — the learned denoiser of the VPL GI pass: a kernel-predicting
 * convolutional network that replaces the a-trous filter of
 * vpl_pbr_gi_denoise.fs for low-sample gathers, with a per-tile fallback to
 * the a-trous filter where the network's own error estimate is high.
 *
 * The reference has no learned denoiser; the network stands beside
 * sgi_denoise (vpl_pbr_gi_denoise.fs:142-435, StratusRendererBackend.cpp:
 * 1879-1951) with the same inputs (G-buffer, gather colour / reservoir, lit
 * screen) and the same outputs (combined = screen + albedo * gi, gi,
 * reservoir, history).
 *
 * Network: a 5-level encoder-decoder of 3x3 convolutions with 32 channels
 * (levels at 1, 1/2, ..., 1/16 resolution; 2x2 mean pooling down, nearest
 * upsampling plus skip addition up; ReLU after every convolution but the
 * filter heads).  A head per level predicts per pixel the 25 logits of a 5x5
 * filter (softmax-normalised) of that level's radiance, and one blend logit;
 * the level-0 head also predicts the log relative squared error.
 *
 *   layer 0: features -> e0           (level 0, one convolution at full resolution)
 *   layer 1: pool(e0) -> c1           (level 1)      layer 2: c1 -> e1
 *   layer 3: pool(e1) -> c3           (level 2)      layer 4: c3 -> e2
 *   layer 5: pool(e2) -> c5           (level 3)      layer 6: c5 -> e3
 *   layer 7: pool(e3) -> c7           (level 4)      layer 8: c7 -> e4
 *   layer 9: e3 + up(e4) -> d3        layer 10: e2 + up(d3) -> d2
 *   layer 11: e1 + up(d2) -> d1
 *   heads (no ReLU, FP32 logits): layer 12: u0 = e0 + up(d1) (level 0),
 *   13: d1, 14: d2, 15: d3, 16: e4 (level 4)
 *
 * Logit channels: 0..24 filter taps (tap j at offset (j % 5 - 2, j / 5 - 2),
 * clamped to the image), SGI_KPCN_ERR_CHANNEL the error estimate (level 0),
 * SGI_KPCN_BLEND_CHANNEL the blend (levels 0..3); the others are unused.
 *
 * Reconstruction, coarse to fine: r_0 = the noisy radiance, r_l = the 2x2
 * mean of r_(l-1) (the radiance pyramid); F_l = the level's filter applied
 * to r_l; o_4 = F_4 and o_l = F_l + sigmoid(blend_l) (up(o_(l+1)) - F_l),
 * where up is bilinear: fine pixel x samples the coarser grid at
 * (x + 0.5) / 2 - 0.5 (weights 1/4, 3/4), indices clamped to the grid.  The
 * output is o_0.
 *
 * Tensors: activations and their gradients are BF16 [h][w][32] (channels
 * innermost, 64 B per pixel); level l has ceil(h / 2^l) x ceil(w / 2^l)
 * pixels.  Weights of one layer are BF16 [9 taps][32 out][32 in] (tap =
 * 3 * dy + dx over the 3x3 window, dy / dx = 0, 1, 2 for offsets -1, 0, +1;
 * zero padding at the borders); the matching FP32 parameters of a layer are
 * [9][32][32] weights followed by 32 biases (SGI_KPCN_LAYER_PARAMS floats).
 * Matrix work runs on MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation.
 *
 * Errors and streams as sgi.h.
 */
#ifndef SGI_KPCN_H
#define SGI_KPCN_H

#include "sgi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SGI_KPCN_CHANNELS 32
#define SGI_KPCN_LEVELS 5
#define SGI_KPCN_LAYERS 17
#define SGI_KPCN_LAYER_WEIGHTS (9 * 32 * 32)
#define SGI_KPCN_LAYER_PARAMS (SGI_KPCN_LAYER_WEIGHTS + 32)
#define SGI_KPCN_PARAMS (SGI_KPCN_LAYERS * SGI_KPCN_LAYER_PARAMS)
#define SGI_KPCN_TAPS 25          /* 5x5 predicted filter */
#define SGI_KPCN_ERR_CHANNEL 25   /* level-0 logit of the log relative squared error */
#define SGI_KPCN_BLEND_CHANNEL 26 /* logit of the blend with the coarser level */
#define SGI_KPCN_FEATURES 10      /* log radiance 3, albedo 3, normal 3, depth 1 */
#define SGI_KPCN_TILE 16          /* fallback decision granularity (pixels) */
#define SGI_KPCN_RADIANCE_SCALE 64.0f  /* radiance feature = log(1 + 64 radiance) */

/* A trained network in inference form.
 *
 * FP8 inference (weights_fp8 != NULL): layers 1..16 run on OCP FP8 E4M3
 * activations (stored 32 B per pixel) and weights with FP32 accumulation;
 * layer 0 reads the BF16 features with weights[0] and stores E4M3. Each
 * activation tensor a carries per-channel power-of-two scales s (stored
 * q = e4m3(min(a s, 448))); a level's skip tensor e_l, the decoder output d_{l+1}
 * added to it (e_4 for level 3) and their sum share one scale, and a pooled
 * tensor keeps its source's. Layer l's FP8 weights are
 * e4m3(W[t][o][i] / s_in[i] * w_o) with a power-of-two w_o per output channel,
 * and its epilogue row [k1[32], k2[32]] gives the output channel o as
 * fma(acc_o, k1_o, k2_o): an activation stores e4m3(min(max(., 0), 448))
 * (k1 = s_out / w_o, k2 = b s_out; layer 0: k1 = s_out), a filter head keeps
 * the FP32 logit (k1 = 1 / w_o, k2 = b). */
typedef struct sgi_kpcn_net {
    const uint16_t* weights;     /* device BF16 [SGI_KPCN_LAYERS][9][32][32] */
    const float* biases;         /* device FP32 [SGI_KPCN_LAYERS][32] */
    const uint8_t* weights_fp8;  /* NULL (BF16 inference) or device E4M3 [SGI_KPCN_LAYERS][9][32][32] */
    const float* epilogue;       /* with weights_fp8: device FP32 [SGI_KPCN_LAYERS][2][32] */
} sgi_kpcn_net;

/* ---- building blocks (training and tests) -------------------------------- */

/* out = act(conv3x3(in (x) [mask > 0]) + bias): in BF16 [h][w][32]; mask NULL
 * or BF16 [h][w][32] (the input is zeroed where mask <= 0: the ReLU gradient
 * gate); weights BF16 [9][32][32]; bias FP32 [32] or NULL; relu 0/1; out BF16
 * [h][w][32] (out_f32 = 0) or FP32 [h][w][32] (out_f32 = 1). */
int sgi_kpcn_conv3x3(const uint16_t* in, const uint16_t* mask, const uint16_t* weights, const float* bias,
                     int32_t relu, int32_t height, int32_t width, void* out, int32_t out_f32, void* stream);

/* FP8 inference layer: out = e4m3(min(max(fma(acc, k1, k2), 0), 448)) per
 * output channel, acc = sum over taps and input channels of E4M3 weights [9][32][32]
 * times E4M3 inputs (FP32 accumulation); in, out E4M3 [h][w][32] (zero outside
 * the image); epilogue FP32 [2][32] = (k1, k2). */
int sgi_kpcn_conv3x3_fp8(const uint8_t* in, const uint8_t* weights, const float* epilogue, int32_t height,
                         int32_t width, uint8_t* out, void* stream);

/* BF16 of n floats (n <= 2^38), rounded to the nearest even: the conversion
 * every BF16 activation, gradient and weight store of the network uses. */
int sgi_kpcn_bf16_encode(const float* in, int64_t n, uint16_t* out, void* stream);

/* E4M3 bytes of n floats: clamped to +-448, rounded to the nearest even E4M3
 * value (OCP FP8; the conversion the FP8 layers store with). */
int sgi_kpcn_fp8_encode(const float* in, int64_t n, uint8_t* out, void* stream);

/* Weight gradient of one layer: grad[t][o][i] = sum_p dz[p][o] * x[p + off(t)][i]
 * and grad[9216 + o] = sum_p dz[p][o], with dz = dy (x) [mask > 0] (mask NULL:
 * dz = dy).  x, dy, mask BF16 [h][w][32]; grad FP32 [SGI_KPCN_LAYER_PARAMS]
 * (overwritten).  partials: device FP32 workspace of
 * sgi_kpcn_wgrad_partials(h, w) floats. */
size_t sgi_kpcn_wgrad_partials(int32_t height, int32_t width);
int sgi_kpcn_conv3x3_wgrad(const uint16_t* x, const uint16_t* dy, const uint16_t* mask, int32_t height,
                           int32_t width, float* partials, float* grad, void* stream);

/* 2x2 mean pooling: out[y][x] = mean of the in-range in[2y + i][2x + j].
 * Backward: din[child] (+)= dout[parent] / (number of in-range children);
 * accumulate 0 overwrites din, 1 adds to it. Sizes are the finer level's. */
int sgi_kpcn_pool(const uint16_t* in, int32_t height, int32_t width, uint16_t* out, void* stream);
int sgi_kpcn_pool_backward(const uint16_t* dout, int32_t height, int32_t width, uint16_t* din,
                           int32_t accumulate, void* stream);

/* out = skip + up(low) (nearest): skip, out BF16 at (height, width), low at
 * the coarser level.  Backward of the low branch: dlow[Y][X] (+)= sum of the
 * in-range dout[2Y + i][2X + j] (the skip branch's gradient is dout itself);
 * accumulate 0 overwrites dlow, 1 adds to it. */
int sgi_kpcn_upsample_add(const uint16_t* skip, const uint16_t* low, int32_t height, int32_t width,
                          uint16_t* out, void* stream);
int sgi_kpcn_upsample_backward(const uint16_t* dout, int32_t height, int32_t width, uint16_t* dlow,
                               int32_t accumulate, void* stream);

/* The radiance pyramid: levels 1..4 of r_l = 2x2 mean (in-range children) of
 * r_(l-1), r_0 = radiance FP32 [h][w][3].  pyramid: FP32, the levels one
 * after the other ([h_1][w_1][3] first, h_l = ceil(h_(l-1) / 2)), of
 * sgi_kpcn_pyramid_floats(h, w) floats. */
size_t sgi_kpcn_pyramid_floats(int32_t height, int32_t width);
int sgi_kpcn_radiance_pyramid(const float* radiance, int32_t height, int32_t width, float* pyramid,
                              void* stream);

/* Network input from a gather: radiance = reservoir.rgb / max(reservoir.a,
 * 1e-5) (FP32 [h][w][3], the a-trous filter's pass-0 normalisation,
 * vpl_pbr_gi_denoise.fs:329) and BF16 features [h][w][32]: log(1 +
 * SGI_KPCN_RADIANCE_SCALE * radiance), albedo, decoded normal, log(1 + view
 * distance) / 8, zeros.
 * Needs sgi_set_frame. */
int sgi_kpcn_features(sgi_ctx* ctx, const sgi_gbuffer* gbuffer, const float* reservoir, uint16_t* features,
                      float* radiance, void* stream);

/* One level of the reconstruction: F = sum_j softmax(logits[0..24])_j *
 * radiance(p + o_j) over the 5x5 window (clamped to the image); out = F when
 * coarse is NULL, else F + sigmoid(logits[SGI_KPCN_BLEND_CHANNEL]) *
 * (up(coarse) - F); err = logits[SGI_KPCN_ERR_CHANNEL].  logits FP32
 * [h][w][32]; radiance FP32 [h][w][3] (this level's); coarse FP32
 * [ceil(h/2)][ceil(w/2)][3] or NULL; out FP32 [h][w][3]; err FP32 [h][w] or
 * NULL. */
int sgi_kpcn_apply(const float* logits, const float* radiance, const float* coarse, int32_t height,
                   int32_t width, float* out, float* err, void* stream);

/* Backward of sgi_kpcn_apply for the output gradient dout (FP32 [h][w][3]):
 * dlogits BF16 [h][w][32] (filter and blend channels; the error channel
 * takes derr (FP32 [h][w]) or 0 when derr is NULL; the rest 0) and, when
 * coarse is given, dup FP32 [h][w][3] = the gradient of up(coarse). */
int sgi_kpcn_apply_backward(const float* logits, const float* radiance, const float* coarse, const float* dout,
                            const float* derr, int32_t height, int32_t width, uint16_t* dlogits, float* dup,
                            void* stream);

/* The transpose of the bilinear up(): dcoarse FP32 [ceil(h/2)][ceil(w/2)][3]
 * = sum over fine pixels of their bilinear weight times dup (FP32 [h][w][3]);
 * overwrites dcoarse. */
int sgi_kpcn_upsample_bilinear_backward(const float* dup, int32_t height, int32_t width, float* dcoarse,
                                        void* stream);

/* Training loss and its gradient on the reconstruction out (FP32 [h][w][3])
 * and the level-0 logits' error channel err:
 * loss = mean_{p,c} (out - t)^2 / (t^2 + 0.01)
 *      + err_weight * mean_p (err - log(r_p + 1e-4))^2,
 * r_p = mean_c (out - t)^2 / (t^2 + 0.01) (held constant).  target FP32
 * [h][w][3]; dout FP32 [h][w][3] and derr FP32 [h][w] receive the gradients
 * times scale (1 / batch); loss_sums FP32 [2] accumulates the two sums
 * (before the means). */
int sgi_kpcn_loss(const float* out, const float* logits, const float* target, int32_t height, int32_t width,
                  float err_weight, float scale, float* dout, float* derr, float* loss_sums, void* stream);

/* sgi_kpcn_loss with the error channel trained on the log ratio of the
 * reconstruction's relative squared error to that of another output of the
 * same input, alternative (FP32 [h][w][3]; the a-trous filter's output):
 * the error term is err_weight * mean_p (err - log(r_p + 1e-4) + log(a_p +
 * 1e-4))^2, a_p = mean_c (alternative - t)^2 / (t^2 + 0.01) (held constant),
 * so that exp(err) estimates (r_p + 1e-4) / (a_p + 1e-4), the quantity whose
 * tile mean sgi_denoise_learned compares with its threshold when the network
 * was trained this way.  The reconstruction term and its gradient are
 * sgi_kpcn_loss's. */
int sgi_kpcn_loss_ratio(const float* out, const float* logits, const float* target, const float* alternative,
                        int32_t height, int32_t width, float err_weight, float scale, float* dout, float* derr,
                        float* loss_sums, void* stream);

/* One Adam step over all SGI_KPCN_PARAMS parameters (FP32 params, grad, m, v)
 * followed by the BF16 copies: weights [layers][9][32][32], weights_t (the
 * transposed, tap-mirrored weights of the input gradient: weights_t[l][t][i][o]
 * = W[l][8 - t][o][i]) and biases FP32 [layers][32]. */
int sgi_kpcn_adam(float* params, const float* grad, float* m, float* v, float lr, float beta1, float beta2,
                  float eps, int32_t step, uint16_t* weights, uint16_t* weights_t, float* biases, void* stream);

/* The BF16 / FP32 copies sgi_kpcn_adam writes after its step, from params
 * alone (initial weights, loaded weights). */
int sgi_kpcn_pack_weights(const float* params, uint16_t* weights, uint16_t* weights_t, float* biases, void* stream);

/* ---- inference ------------------------------------------------------------ */

/* Device workspace of sgi_denoise_learned at a viewport (bytes). */
size_t sgi_kpcn_workspace_bytes(int32_t height, int32_t width);

/* The learned denoise of one frame: features -> network -> multi-scale
 * reconstruction of the noisy radiance; then per 16x16 tile, if the mean of
 * exp(err) over its pixels exceeds `threshold`, the tile takes the a-trous
 * result instead.  exp(err) is the network's predicted relative squared error
 * when its error channel was trained by sgi_kpcn_loss, and the predicted ratio
 * of that error to the a-trous filter's when trained by sgi_kpcn_loss_ratio
 * (the shipped weights; threshold 1.2).  The a-trous passes run only over
 * the flagged tiles and a 28-pixel margin, and not at all when no tile is
 * flagged.  Outputs as
 * sgi_denoise (single frame: history = 1, reservoirOut = (gi, 1)); prevGi /
 * prevHistory as sgi_denoise.  fallbackTiles (device int32 [1 + tiles], may be
 * NULL) receives the number of fallback tiles and one flag per tile.
 * threshold <= 0 disables the network's output (every tile falls back);
 * a threshold of +inf disables the fallback.  workspace: device memory of
 * sgi_kpcn_workspace_bytes(); net: trained weights. */
int sgi_denoise_learned(sgi_ctx* ctx, const sgi_gbuffer* cur, const sgi_gbuffer* prev,
                        const int32_t* visibleIndices, const float* screen, const float* color,
                        const float* reservoir, const float* prevGi, const float* prevHistory,
                        const sgi_kpcn_net* net, float threshold, void* workspace, size_t workspaceBytes,
                        float* combined, float* giOut, float* reservoirOut, float* historyOut,
                        int32_t* fallbackTiles, void* stream);

/* The learned form of a batch of views (sgi_views_*): VPL culling, atlas
 * entries and direct light stay one launch over every view; then per view
 * the sampled gather (sgi_gather_ex SAMPLES: `samples` hashed draws per pixel
 * over every visible VPL, rejectionLimit as sgi_gather_opts, FP32 shading)
 * and sgi_denoise_learned with `net` and `threshold` (no previous G-buffer:
 * prevGi / prevHistory of the view's sgi_view_io), each view's outputs equal
 * to that chain run on the view alone.  Tile lists are not computed (the
 * views' tiles buffers are left as they are).  samples = 0 restores the
 * tile-list form.  Needs sgi_views_set first: the network's workspace is
 * sized for the batch's largest view (a later sgi_views_set with larger
 * views grows it); net and the weights it names must stay valid while the
 * batch runs or replays.  A captured graph keeps the form it was captured
 * with: sgi_views_replay refuses a batch whose form changed since. */
int sgi_views_set_learned(sgi_views* views, int32_t samples, int32_t rejectionLimit, const sgi_kpcn_net* net,
                          float threshold);

#ifdef __cplusplus
}
#endif

#endif /* SGI_KPCN_H */
