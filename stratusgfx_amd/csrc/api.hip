// C ABI of the GI pass: context lifetime, argument validation, and the
// per-frame orchestration that mirrors RendererBackend's GI steps
// (StratusRendererBackend.cpp:1982, 2010-2045).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <new>
#include <string>

#include "../../include/sgi.h"
#include "../../include/sgi_kpcn.h"
#include "launch.h"

void sgi_set_error(const std::string& msg);

struct sgi_ctx {
    sgi_config cfg;
    int tilesX = 0, tilesY = 0;
    bool hasFrame = false, hasCsm = false;
    sgi_frame frame{};
    sgi_dev::FrameArgs frameArgs{};
    sgi_dev::CsmArgs csmArgs{};
    sgi_launch::DenoiseBuffers ws{};
    uint8_t* visFlags = nullptr;  // [SGI_MAX_TOTAL_VPLS_BEFORE_CULLING] per-VPL sun visibility
    void* vplPrep = nullptr;      // [SGI_MAX_TOTAL_VPLS_PER_FRAME] packed records of the sampled gathers
    int32_t* cubeLists = nullptr; // per (light, face) triangle lists of sgi_render_vpl_cubes, grown on demand
    size_t cubeListInts = 0;
    float* csmTris = nullptr;     // per cascade transformed triangle lists of sgi_render_csm, grown on demand
    size_t csmTrisFloats = 0;
    int device = 0;
    hipEvent_t screenReady = nullptr;  // sgi_set_screen_event: waited for before the passes that read screen
};

namespace {

int fail(int code, const std::string& msg) {
    sgi_set_error(msg);
    return code;
}

int hipFail(hipError_t e, const char* what) {
    return fail(SGI_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int checkGbuffer(const sgi_ctx* c, const sgi_gbuffer* g, const char* who, bool needMaterial) {
    if (!g) return fail(SGI_E_INVALID, std::string(who) + ": NULL gbuffer");
    if (g->width != c->cfg.width || g->height != c->cfg.height)
        return fail(SGI_E_INVALID, std::string(who) + ": gbuffer size differs from the context's viewport");
    if (g->packed) {
        const sgi_gbuffer_packed& k = *g->packed;
        if (k.width != g->width || k.height != g->height)
            return fail(SGI_E_INVALID, std::string(who) + ": packed attachments differ in size from the gbuffer");
        if (!g->depth || !k.normals) return fail(SGI_E_INVALID, std::string(who) + ": gbuffer depth / normals are NULL");
        if (needMaterial && (!k.albedo || !k.baseReflectivity || !k.roughnessMetallicAmbient))
            return fail(SGI_E_INVALID, std::string(who) + ": packed albedo / reflectivity / roughness is NULL");
        return SGI_OK;
    }
    if (!g->depth || !g->normal) return fail(SGI_E_INVALID, std::string(who) + ": gbuffer depth / normal is NULL");
    if (needMaterial && (!g->albedo || !g->baseReflectivity || !g->roughnessMetallicAmbient))
        return fail(SGI_E_INVALID, std::string(who) + ": gbuffer albedo / reflectivity / roughness plane is NULL");
    return SGI_OK;
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// ---- scratch of sgi_gather_samples_noalloc ---------------------------------
// The packed visible-VPL records the sampled gather reads with their binary16 positions
// (sgi_launch::packedVplBytes() = 64 + 8 B each), for up to 4480 visible VPLs: 315 KB of device
// memory per GPU, held until the process ends.  In host code `scratch` names a host-side stand-in, so its device address is looked up
// once per device (hipGetSymbolAddress, which allocates nothing) and only that address is used.
constexpr int kNoAllocMaxVisible = SGI_MAX_TOTAL_VPLS_PER_FRAME;
constexpr int kNoAllocRecordBytes = 72;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocMaxVisible * kNoAllocRecordBytes];

std::once_flag g_scratchOnce[kMaxDevices];
void* g_scratchPtr[kMaxDevices];
hipError_t g_scratchErr[kMaxDevices];

int noAllocScratch(void** out) {
    if (sgi_launch::packedVplBytes() != (size_t)kNoAllocRecordBytes)
        return fail(SGI_E_STATE, "sgi_gather_samples_noalloc: packed record size differs from the scratch layout");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hipFail(e, "sgi_gather_samples_noalloc: hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(SGI_E_INVALID, "sgi_gather_samples_noalloc: device index >= 64");
    std::call_once(g_scratchOnce[dev], [dev] {
        g_scratchErr[dev] = hipGetSymbolAddress(&g_scratchPtr[dev], HIP_SYMBOL(scratch));
    });
    if (g_scratchErr[dev] != hipSuccess)
        return hipFail(g_scratchErr[dev], "sgi_gather_samples_noalloc: hipGetSymbolAddress(scratch)");
    *out = g_scratchPtr[dev];
    return SGI_OK;
}

}  // namespace

extern "C" {

int sgi_create(const sgi_config* config, sgi_ctx** out) {
    if (!config || !out) return fail(SGI_E_INVALID, "sgi_create: NULL argument");
    *out = nullptr;
    const sgi_config& c = *config;
    if (c.width <= 0 || c.height <= 0 || c.width > 16384 || c.height > 16384)
        return fail(SGI_E_INVALID, "sgi_create: viewport must be within 1..16384 on each side");
    if (c.tile_size != 8 && c.tile_size != 16 && c.tile_size != 32)
        return fail(SGI_E_INVALID, "sgi_create: tile_size must be 8, 16 or 32");
    if (c.max_vpls_per_tile < 1 || c.max_vpls_per_tile > SGI_MAX_VPLS_PER_TILE_LIMIT)
        return fail(SGI_E_INVALID, "sgi_create: max_vpls_per_tile must be within 1..64");
    sgi_ctx* ctx = new (std::nothrow) sgi_ctx();
    if (!ctx) return fail(SGI_E_INVALID, "sgi_create: out of host memory");
    ctx->cfg = c;
    ctx->tilesX = (c.width + c.tile_size - 1) / c.tile_size;
    ctx->tilesY = (c.height + c.tile_size - 1) / c.tile_size;
    hipError_t e = hipGetDevice(&ctx->device);
    const size_t plane = (size_t)c.width * c.height * sizeof(float4);
    if (e == hipSuccess) e = hipMalloc(&ctx->ws.resA, plane);
    if (e == hipSuccess) e = hipMalloc(&ctx->ws.resB, plane);
    if (e == hipSuccess) e = hipMalloc(&ctx->ws.normals, plane);
    if (e == hipSuccess) e = hipMalloc(&ctx->visFlags, SGI_MAX_TOTAL_VPLS_BEFORE_CULLING);
    if (e == hipSuccess) e = hipMalloc(&ctx->vplPrep, SGI_MAX_TOTAL_VPLS_PER_FRAME * sgi_launch::packedVplBytes());
    if (e != hipSuccess) {
        sgi_destroy(ctx);
        return hipFail(e, "sgi_create: workspace allocation");
    }
    *out = ctx;
    return SGI_OK;
}

void sgi_destroy(sgi_ctx* ctx) {
    if (!ctx) return;
    if (ctx->ws.resA) (void)hipFree(ctx->ws.resA);
    if (ctx->ws.resB) (void)hipFree(ctx->ws.resB);
    if (ctx->ws.normals) (void)hipFree(ctx->ws.normals);
    if (ctx->visFlags) (void)hipFree(ctx->visFlags);
    if (ctx->vplPrep) (void)hipFree(ctx->vplPrep);
    if (ctx->cubeLists) (void)hipFree(ctx->cubeLists);
    if (ctx->csmTris) (void)hipFree(ctx->csmTris);
    delete ctx;
}

int sgi_tile_counts(const sgi_ctx* ctx, int32_t* tx, int32_t* ty) {
    if (!ctx || !tx || !ty) return fail(SGI_E_INVALID, "sgi_tile_counts: NULL argument");
    *tx = ctx->tilesX;
    *ty = ctx->tilesY;
    return SGI_OK;
}

int sgi_set_frame(sgi_ctx* ctx, const sgi_frame* f) {
    if (!ctx || !f) return fail(SGI_E_INVALID, "sgi_set_frame: NULL argument");
    if (!(f->framesPerSecond > 0.0f)) return fail(SGI_E_INVALID, "sgi_set_frame: framesPerSecond must be > 0");
    ctx->frame = *f;
    ctx->frameArgs = sgi_dev::frameArgs(*f);
    ctx->hasFrame = true;
    return SGI_OK;
}

int sgi_set_csm(sgi_ctx* ctx, const sgi_csm* c) {
    if (!ctx || !c) return fail(SGI_E_INVALID, "sgi_set_csm: NULL argument");
    if (c->resolution <= 0 || c->resolution > 16384 || !c->depth)
        return fail(SGI_E_INVALID, "sgi_set_csm: resolution must be within 1..16384 and depth non-NULL");
    sgi_dev::CsmArgs& a = ctx->csmArgs;
    a.depth = c->depth;
    a.res = c->resolution;
    for (int i = 0; i < SGI_NUM_CASCADES; ++i)
        for (int k = 0; k < 16; ++k) a.projViews[i][k] = c->cascadeProjViews[i][k];
    for (int i = 0; i < SGI_NUM_CASCADES - 1; ++i)
        for (int k = 0; k < 4; ++k) a.planes[i][k] = c->cascadePlanes[i][k];
    // pbr.glsl:116: infiniteLightDepthBias / (infiniteLightZfar - infiniteLightZnear)
    a.bias = c->depthBias / (c->zfar - c->znear);
    ctx->hasCsm = true;
    return SGI_OK;
}

int sgi_taa(sgi_ctx* ctx, const float* screen, const float* prevScreen, const float* velocity, float* out,
            void* stream) {
    if (!ctx) return fail(SGI_E_INVALID, "sgi_taa: NULL context");
    if (!screen || !prevScreen || !out) return fail(SGI_E_INVALID, "sgi_taa: NULL buffer");
    if (out == screen || out == prevScreen) return fail(SGI_E_INVALID, "sgi_taa: out must not alias an input");
    const hipError_t e =
        sgi_launch::taa(screen, prevScreen, velocity, ctx->cfg.width, ctx->cfg.height, out, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_taa");
}

int sgi_taa_history(sgi_ctx* ctx, const float* taaOut, float* prevScreen, void* stream) {
    if (!ctx) return fail(SGI_E_INVALID, "sgi_taa_history: NULL context");
    if (!taaOut || !prevScreen) return fail(SGI_E_INVALID, "sgi_taa_history: NULL buffer");
    const hipError_t e =
        sgi_launch::taaHistory(taaOut, (long)ctx->cfg.width * ctx->cfg.height, prevScreen, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_taa_history");
}

int sgi_fxaa(sgi_ctx* ctx, const uint32_t* tonemapped, const sgi_fxaa_params* params, uint32_t* out, void* stream) {
    if (!ctx) return fail(SGI_E_INVALID, "sgi_fxaa: NULL context");
    if (!tonemapped || !out) return fail(SGI_E_INVALID, "sgi_fxaa: NULL buffer");
    if (out == tonemapped) return fail(SGI_E_INVALID, "sgi_fxaa: out must not alias the input");
    if (((uintptr_t)tonemapped & 3) || ((uintptr_t)out & 3)) return fail(SGI_E_INVALID, "sgi_fxaa: misaligned buffer");
    sgi_fxaa_params p{0.0625f, 0.125f, 1.0f};  // fxaa_smoothing.fs:77, 90, 101
    if (params) p = *params;
    if (!std::isfinite(p.contrastThreshold) || !std::isfinite(p.relativeThreshold) ||
        !std::isfinite(p.subpixelBlending))
        return fail(SGI_E_INVALID, "sgi_fxaa: parameters must be finite");
    const hipError_t e = sgi_launch::fxaa(tonemapped, ctx->cfg.width, ctx->cfg.height, p.contrastThreshold,
                                          p.relativeThreshold, p.subpixelBlending, out, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_fxaa");
}

int sgi_tonemap(sgi_ctx* ctx, const float* lit, float gamma, uint32_t* out, void* stream) {
    if (!ctx) return fail(SGI_E_INVALID, "sgi_tonemap: NULL context");
    if (!lit || !out) return fail(SGI_E_INVALID, "sgi_tonemap: NULL buffer");
    if (!(gamma > 0.0f) || !std::isfinite(gamma)) return fail(SGI_E_INVALID, "sgi_tonemap: gamma must be > 0");
    if (((uintptr_t)lit & 3) || ((uintptr_t)out & 3)) return fail(SGI_E_INVALID, "sgi_tonemap: misaligned buffer");
    const hipError_t e =
        sgi_launch::tonemap(lit, (long)ctx->cfg.width * ctx->cfg.height, gamma, out, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_tonemap");
}

int sgi_gbuffer_pack(sgi_ctx* ctx, const sgi_gbuffer* planes, const sgi_gbuffer_packed* out, void* stream) {
    if (!ctx || !planes || !out) return fail(SGI_E_INVALID, "sgi_gbuffer_pack: NULL argument");
    if (planes->width != ctx->cfg.width || planes->height != ctx->cfg.height || out->width != planes->width ||
        out->height != planes->height)
        return fail(SGI_E_INVALID, "sgi_gbuffer_pack: size differs from the context's viewport");
    if (!planes->normal || !planes->albedo || !planes->baseReflectivity || !planes->roughnessMetallicAmbient ||
        !out->normals || !out->albedo || !out->baseReflectivity || !out->roughnessMetallicAmbient)
        return fail(SGI_E_INVALID, "sgi_gbuffer_pack: NULL plane or attachment");
    const auto al = [](const void* p, uintptr_t a) { return ((uintptr_t)p & (a - 1)) == 0; };
    if (!al(out->normals, 4) || !al(out->albedo, 4) || !al(out->roughnessMetallicAmbient, 4) ||
        !al(out->baseReflectivity, 2) || (out->velocity && !al(out->velocity, 4)) || !al(planes->normal, 4) ||
        !al(planes->albedo, 4) || !al(planes->baseReflectivity, 4) || !al(planes->roughnessMetallicAmbient, 4) ||
        (planes->velocity && !al(planes->velocity, 4)))
        return fail(SGI_E_INVALID, "sgi_gbuffer_pack: misaligned plane (texel alignment)");
    const hipError_t e = sgi_launch::gbufferPack(*planes, *out, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_gbuffer_pack");
}

int sgi_gbuffer_unpack(sgi_ctx* ctx, const sgi_gbuffer_packed* g, float* normal, float* albedo,
                       float* baseReflectivity, float* roughnessMetallicAmbient, float* velocity, void* stream) {
    if (!ctx || !g) return fail(SGI_E_INVALID, "sgi_gbuffer_unpack: NULL argument");
    if (g->width != ctx->cfg.width || g->height != ctx->cfg.height)
        return fail(SGI_E_INVALID, "sgi_gbuffer_unpack: size differs from the context's viewport");
    if (!g->normals || !g->albedo || !g->baseReflectivity || !g->roughnessMetallicAmbient || !normal || !albedo ||
        !baseReflectivity || !roughnessMetallicAmbient || (g->velocity && !velocity))
        return fail(SGI_E_INVALID, "sgi_gbuffer_unpack: NULL plane");
    const auto al = [](const void* p, uintptr_t a) { return ((uintptr_t)p & (a - 1)) == 0; };
    if (!al(g->normals, 4) || !al(g->albedo, 4) || !al(g->roughnessMetallicAmbient, 4) ||
        !al(g->baseReflectivity, 2) || (g->velocity && !al(g->velocity, 4)) || !al(normal, 4) || !al(albedo, 4) ||
        !al(baseReflectivity, 4) || !al(roughnessMetallicAmbient, 4) || (velocity && !al(velocity, 4)))
        return fail(SGI_E_INVALID, "sgi_gbuffer_unpack: misaligned plane (texel alignment)");
    const hipError_t e = sgi_launch::gbufferUnpack(*g, normal, albedo, baseReflectivity, roughnessMetallicAmbient,
                                                   g->velocity ? velocity : nullptr, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_gbuffer_unpack");
}

int sgi_cull_global(sgi_ctx* ctx, const sgi_vpl* vplData, int32_t n, sgi_vpl* vplUpdatedData, int32_t* visibleIndices,
                    void* stream) {
    if (!ctx) return fail(SGI_E_INVALID, "sgi_cull_global: NULL context");
    if (!ctx->hasFrame || !ctx->hasCsm) return fail(SGI_E_STATE, "sgi_cull_global: sgi_set_frame and sgi_set_csm first");
    if (n < 0 || n > SGI_MAX_TOTAL_VPLS_BEFORE_CULLING)
        return fail(SGI_E_INVALID, "sgi_cull_global: n must be within 0..10000");
    if (!visibleIndices || (n > 0 && (!vplData || !vplUpdatedData)))
        return fail(SGI_E_INVALID, "sgi_cull_global: NULL buffer");
    if (n > 0 && (!aligned16(vplData) || !aligned16(vplUpdatedData)))
        return fail(SGI_E_INVALID, "sgi_cull_global: VPL records must be 16-byte aligned");
    const hipError_t e = sgi_launch::cullGlobal(ctx->csmArgs, ctx->frameArgs, vplData, n, vplUpdatedData,
                                                visibleIndices, ctx->visFlags, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_cull_global");
}

int sgi_visible_atlas_entries(sgi_ctx* ctx, const sgi_atlas_entry* shadowIndices, const int32_t* visibleIndices,
                              int32_t maxCount, sgi_atlas_entry* out, void* stream) {
    if (!ctx) return fail(SGI_E_INVALID, "sgi_visible_atlas_entries: NULL context");
    if (maxCount < 0 || maxCount > SGI_MAX_TOTAL_VPLS_PER_FRAME)
        return fail(SGI_E_INVALID, "sgi_visible_atlas_entries: maxCount must be within 0..4480");
    if (!shadowIndices || !visibleIndices || (maxCount > 0 && !out))
        return fail(SGI_E_INVALID, "sgi_visible_atlas_entries: NULL buffer");
    const hipError_t e =
        sgi_launch::gatherVisibleEntries(shadowIndices, visibleIndices, maxCount, out, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_visible_atlas_entries");
}

int sgi_cull_tiles(sgi_ctx* ctx, const sgi_gbuffer* g, const sgi_vpl* vpls, const int32_t* visibleIndices,
                   int32_t remap, sgi_tile_stage1* stage1, int32_t* tiles, void* stream) {
    if (!ctx) return fail(SGI_E_INVALID, "sgi_cull_tiles: NULL context");
    if (!ctx->hasFrame) return fail(SGI_E_STATE, "sgi_cull_tiles: sgi_set_frame first");
    int r = checkGbuffer(ctx, g, "sgi_cull_tiles", false);
    if (r != SGI_OK) return r;
    if (!vpls || !visibleIndices || !tiles) return fail(SGI_E_INVALID, "sgi_cull_tiles: NULL buffer");
    const hipError_t e = sgi_launch::cullTiles(sgi_dev::gbufArgs(*g), ctx->frameArgs, ctx->cfg.tile_size, ctx->tilesX,
                                               ctx->tilesY, ctx->cfg.max_vpls_per_tile, vpls, visibleIndices, remap,
                                               stage1, tiles, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_cull_tiles");
}

namespace {

int cubeArgs(const sgi_cube_atlases* cubes, const char* who, sgi_dev::CubeArgs& ca) {
    if (cubes->resolution <= 0 || cubes->resolution > 4096 ||
        (cubes->format != SGI_TEXEL_F32 && cubes->format != SGI_TEXEL_F16 && cubes->format != SGI_TEXEL_UNORM16) ||
        cubes->layers_per_atlas <= 0 ||
        cubes->layers_per_atlas > SGI_MAX_TOTAL_SHADOWS_PER_ATLAS || cubes->num_atlases <= 0 ||
        cubes->num_atlases > SGI_MAX_TOTAL_SHADOW_ATLASES)
        return fail(SGI_E_INVALID, std::string(who) + ": bad cube atlas description");
    ca = sgi_dev::CubeArgs{};
    for (int a = 0; a < cubes->num_atlases; ++a) {
        if (!cubes->atlas[a]) return fail(SGI_E_INVALID, std::string(who) + ": NULL cube atlas");
        ca.atlas[a] = cubes->atlas[a];
    }
    for (int a = cubes->num_atlases; a < SGI_MAX_TOTAL_SHADOW_ATLASES; ++a) ca.atlas[a] = cubes->atlas[0];
    ca.numAtlases = cubes->num_atlases;
    ca.res = cubes->resolution;
    ca.layers = cubes->layers_per_atlas;
    ca.layerStride = 6u * (uint32_t)cubes->resolution * (uint32_t)cubes->resolution;
    return SGI_OK;
}

int checkGatherArgs(sgi_ctx* ctx, const sgi_gbuffer* g, const sgi_vpl* vpls, const int32_t* visibleIndices,
                    const sgi_atlas_entry* shadowIndices, const sgi_cube_atlases* cubes, float* color,
                    float* reservoir, const char* who) {
    if (!ctx) return fail(SGI_E_INVALID, std::string(who) + ": NULL context");
    if (!ctx->hasFrame) return fail(SGI_E_STATE, std::string(who) + ": sgi_set_frame first");
    int r = checkGbuffer(ctx, g, who, true);
    if (r != SGI_OK) return r;
    if (!vpls || !visibleIndices || !shadowIndices || !cubes || !color || !reservoir)
        return fail(SGI_E_INVALID, std::string(who) + ": NULL buffer");
    if (!aligned16(reservoir)) return fail(SGI_E_INVALID, std::string(who) + ": reservoir must be 16-byte aligned");
    return SGI_OK;
}

}  // namespace

int sgi_color_vpls(sgi_ctx* ctx, const int32_t* visibleIndices, const sgi_atlas_entry* visibleEntries,
                   const sgi_cube_atlases* diffuse, const sgi_cube_atlases* distance, float colorMultiplier,
                   int32_t maxCount, sgi_vpl* vplUpdatedData, void* stream) {
    if (!ctx) return fail(SGI_E_INVALID, "sgi_color_vpls: NULL context");
    if (!ctx->hasFrame) return fail(SGI_E_STATE, "sgi_color_vpls: sgi_set_frame first");
    if (maxCount < 0 || maxCount > SGI_MAX_TOTAL_VPLS_PER_FRAME)
        return fail(SGI_E_INVALID, "sgi_color_vpls: maxCount must be within 0..4480");
    if (!visibleIndices || !diffuse || !distance || (maxCount > 0 && (!visibleEntries || !vplUpdatedData)))
        return fail(SGI_E_INVALID, "sgi_color_vpls: NULL buffer");
    if (diffuse->format != SGI_TEXEL_RGBA8)
        return fail(SGI_E_INVALID, "sgi_color_vpls: the diffuse atlases must be SGI_TEXEL_RGBA8");
    if (diffuse->resolution != distance->resolution || diffuse->layers_per_atlas != distance->layers_per_atlas ||
        diffuse->num_atlases != distance->num_atlases)
        return fail(SGI_E_INVALID, "sgi_color_vpls: the diffuse and distance atlases must share one geometry");
    if (!std::isfinite(colorMultiplier)) return fail(SGI_E_INVALID, "sgi_color_vpls: colorMultiplier must be finite");
    sgi_dev::CubeArgs za, da;
    int r = cubeArgs(distance, "sgi_color_vpls", za);
    if (r != SGI_OK) return r;
    sgi_cube_atlases geometry = *diffuse;  // the geometry checks of a distance atlas
    geometry.format = SGI_TEXEL_F32;
    if ((r = cubeArgs(&geometry, "sgi_color_vpls", da)) != SGI_OK) return r;
    const hipError_t e = sgi_launch::colorVpls(da, za, distance->format, ctx->frameArgs, colorMultiplier,
                                               visibleIndices, maxCount, visibleEntries, vplUpdatedData,
                                               (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_color_vpls");
}

int sgi_gather(sgi_ctx* ctx, const sgi_gbuffer* g, const sgi_vpl* vpls, const int32_t* visibleIndices,
               const sgi_atlas_entry* shadowIndices, const sgi_cube_atlases* cubes, const int32_t* tiles,
               int32_t precision, float* color, float* reservoir, void* stream) {
    int r = checkGatherArgs(ctx, g, vpls, visibleIndices, shadowIndices, cubes, color, reservoir, "sgi_gather");
    if (r != SGI_OK) return r;
    if (precision != SGI_PRECISION_FP32 && precision != SGI_PRECISION_FP16)
        return fail(SGI_E_INVALID, "sgi_gather: unknown precision");
    sgi_dev::CubeArgs ca;
    if ((r = cubeArgs(cubes, "sgi_gather", ca)) != SGI_OK) return r;
    const hipError_t e = sgi_launch::gather(sgi_dev::gbufArgs(*g), ctx->frameArgs, ca, cubes->format,
                                            ctx->cfg.tile_size, ctx->tilesX, ctx->tilesY, ctx->cfg.max_vpls_per_tile,
                                            vpls, visibleIndices, shadowIndices, tiles, precision, ctx->vplPrep, color,
                                            reservoir, (hipStream_t)stream);
    if (e == hipErrorNotSupported) return fail(SGI_E_INVALID, "sgi_gather: this precision is not built");
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_gather");
}

int sgi_gather_ex(sgi_ctx* ctx, const sgi_gbuffer* g, const sgi_vpl* vpls, const int32_t* visibleIndices,
                  const sgi_atlas_entry* shadowIndices, const sgi_cube_atlases* cubes, const int32_t* tiles,
                  const sgi_gather_opts* opts, float* color, float* reservoir, void* stream) {
    int r = checkGatherArgs(ctx, g, vpls, visibleIndices, shadowIndices, cubes, color, reservoir, "sgi_gather_ex");
    if (r != SGI_OK) return r;
    if (!opts) return fail(SGI_E_INVALID, "sgi_gather_ex: NULL options");
    if (opts->mode == SGI_GATHER_REFERENCE || opts->mode == SGI_GATHER_TILES) {
        if (opts->mode == SGI_GATHER_TILES && !tiles) return fail(SGI_E_INVALID, "sgi_gather_ex: tile mode without tiles");
        return sgi_gather(ctx, g, vpls, visibleIndices, shadowIndices, cubes,
                          opts->mode == SGI_GATHER_TILES ? tiles : nullptr, opts->precision, color, reservoir, stream);
    }
    if (opts->mode != SGI_GATHER_ALL && opts->mode != SGI_GATHER_SAMPLES)
        return fail(SGI_E_INVALID, "sgi_gather_ex: unknown mode");
    if (opts->mode == SGI_GATHER_SAMPLES && (opts->samples < 1 || opts->samples > 4096))
        return fail(SGI_E_INVALID, "sgi_gather_ex: samples must be within 1..4096");
    const bool half = opts->precision == SGI_PRECISION_FP16;
    if (opts->precision != SGI_PRECISION_FP32 && !half)
        return fail(SGI_E_INVALID, "sgi_gather_ex: unknown precision");
    if (half && !(opts->mode == SGI_GATHER_SAMPLES && opts->rejection_limit < 0))
        return fail(SGI_E_INVALID, "sgi_gather_ex: FP16 shading is for the fixed-count SAMPLES mode "
                                   "(rejection_limit < 0); SGI_GATHER_ALL shades in FP32");
    sgi_dev::CubeArgs ca;
    if ((r = cubeArgs(cubes, "sgi_gather_ex", ca)) != SGI_OK) return r;
    const hipError_t e = sgi_launch::gatherEx(sgi_dev::gbufArgs(*g), ctx->frameArgs, ca, cubes->format, opts->mode,
                                              opts->samples, opts->rejection_limit, opts->accumulate_fp64,
                                              half ? 1 : 0, vpls,
                                              visibleIndices,
                                              shadowIndices, ctx->vplPrep, color, reservoir, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_gather_ex");
}

int sgi_direct_light(sgi_ctx* ctx, const sgi_gbuffer* g, float* lit, void* stream) {
    if (!ctx) return fail(SGI_E_INVALID, "sgi_direct_light: NULL context");
    if (!ctx->hasFrame || !ctx->hasCsm) return fail(SGI_E_STATE, "sgi_direct_light: sgi_set_frame and sgi_set_csm first");
    int r = checkGbuffer(ctx, g, "sgi_direct_light", true);
    if (r != SGI_OK) return r;
    if (!lit) return fail(SGI_E_INVALID, "sgi_direct_light: NULL output");
    const hipError_t e =
        sgi_launch::directLight(sgi_dev::gbufArgs(*g), ctx->frameArgs, ctx->csmArgs, lit, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_direct_light");
}

int sgi_denoise(sgi_ctx* ctx, const sgi_gbuffer* cur, const sgi_gbuffer* prev, const int32_t* visibleIndices,
                const float* screen, const float* color, const float* reservoir, const float* prevGi,
                const float* prevHistory, float* combined, float* giOut, float* reservoirOut, float* historyOut,
                void* stream) {
    if (!ctx) return fail(SGI_E_INVALID, "sgi_denoise: NULL context");
    if (!ctx->hasFrame) return fail(SGI_E_STATE, "sgi_denoise: sgi_set_frame first");
    int r = checkGbuffer(ctx, cur, "sgi_denoise", false);
    if (r != SGI_OK) return r;
    if (prev && (r = checkGbuffer(ctx, prev, "sgi_denoise(prev)", false)) != SGI_OK) return r;
    if (!visibleIndices || !screen || !color || !reservoir || !combined || !giOut || !reservoirOut || !historyOut)
        return fail(SGI_E_INVALID, "sgi_denoise: NULL buffer");
    if (!aligned16(reservoir) || !aligned16(reservoirOut))
        return fail(SGI_E_INVALID, "sgi_denoise: reservoirs must be 16-byte aligned");
    const sgi_dev::GbufArgs c = sgi_dev::gbufArgs(*cur);
    const sgi_dev::GbufArgs p = sgi_dev::gbufArgs(prev ? *prev : *cur);
    const hipStream_t s = (hipStream_t)stream;
    // passes 0-3 read no direct light; the temporal pass (the combine) is the one that waits for it
    hipError_t e = sgi_launch::denoiseSpatial(c, ctx->frameArgs, visibleIndices, reservoir, reservoirOut, ctx->ws,
                                              nullptr, false, s);
    if (e == hipSuccess && ctx->screenReady) e = hipStreamWaitEvent(s, ctx->screenReady, 0);
    if (e == hipSuccess)
        e = sgi_launch::denoiseTemporal(c, p, ctx->frameArgs, screen, color, reservoirOut, prevGi, prevHistory,
                                        combined, giOut, historyOut, ctx->ws, s);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_denoise");
}

int sgi_set_screen_event(sgi_ctx* ctx, void* event) {
    if (!ctx) return fail(SGI_E_INVALID, "sgi_set_screen_event: NULL context");
    ctx->screenReady = (hipEvent_t)event;
    return SGI_OK;
}

// RenderScene's GI steps without the reference's host readbacks: the visible
// list stays on the device, and the per-slot atlas entries are gathered on the
// device instead of GetOrAllocateShadowMapForLight_ on the host.
int sgi_render_frame(sgi_ctx* ctx, const sgi_frame_io* io, void* stream) {
    if (!ctx || !io) return fail(SGI_E_INVALID, "sgi_render_frame: NULL argument");
    if (io->numVpls < 0 || io->numVpls > SGI_MAX_TOTAL_VPLS_BEFORE_CULLING)
        return fail(SGI_E_INVALID, "sgi_render_frame: numVpls must be within 0..10000");
    if (!io->shadowIndices || !io->visibleShadowIndices)
        return fail(SGI_E_INVALID, "sgi_render_frame: NULL atlas entry buffer");
    if (io->useTiles && !io->tiles) return fail(SGI_E_INVALID, "sgi_render_frame: tile lists requested, tiles NULL");
    // UpdatePointLights_ -> PerformVirtualPointLightCullingStage1_
    int r = sgi_cull_global(ctx, io->vplData, io->numVpls, io->vplUpdatedData, io->visibleIndices, stream);
    if (r != SGI_OK) return r;
    // PerformVirtualPointLightCullingStage2_: atlas entries of the visible slots
    const int maxCount = io->numVpls < SGI_MAX_TOTAL_VPLS_PER_FRAME ? io->numVpls : SGI_MAX_TOTAL_VPLS_PER_FRAME;
    if ((r = sgi_visible_atlas_entries(ctx, io->shadowIndices, io->visibleIndices, maxCount,
                                       io->visibleShadowIndices, stream)) != SGI_OK)
        return r;
    // PerformVirtualPointLightCullingStage2_: the vplColoring dispatch (vpl_light_color.cs)
    if (io->diffuseCubes &&
        (r = sgi_color_vpls(ctx, io->visibleIndices, io->visibleShadowIndices, io->diffuseCubes, io->cubes,
                            io->vplColorMultiplier, maxCount, io->vplUpdatedData, stream)) != SGI_OK)
        return r;
    // deferred lighting pass (pbr.fs)
    if ((r = sgi_direct_light(ctx, io->gbuffer, io->direct, stream)) != SGI_OK) return r;
    // tile culling (the commented-out dispatches, StratusRendererBackend.cpp:1733-1783)
    if (io->useTiles &&
        (r = sgi_cull_tiles(ctx, io->gbuffer, io->vplUpdatedData, io->visibleIndices, 0, nullptr, io->tiles, stream)) !=
            SGI_OK)
        return r;
    // ComputeVirtualPointLightGlobalIllumination_
    if ((r = sgi_gather(ctx, io->gbuffer, io->vplUpdatedData, io->visibleIndices, io->visibleShadowIndices, io->cubes,
                        io->useTiles ? io->tiles : nullptr, io->precision, io->color, io->reservoir, stream)) != SGI_OK)
        return r;
    return sgi_denoise(ctx, io->gbuffer, io->prevGbuffer, io->visibleIndices, io->direct, io->color, io->reservoir,
                       io->prevGi, io->prevHistory, io->combined, io->giOut, io->reservoirOut, io->historyOut, stream);
}

/* ---- learned denoiser ------------------------------------------------------ */

static int dims(int32_t h, int32_t w, const char* who) {
    if (h <= 0 || w <= 0 || h > 16384 || w > 16384)
        return fail(SGI_E_INVALID, std::string(who) + ": size must be within 1..16384 on each side");
    return SGI_OK;
}

int sgi_kpcn_conv3x3(const uint16_t* in, const uint16_t* mask, const uint16_t* weights, const float* bias,
                     int32_t relu, int32_t height, int32_t width, void* out, int32_t out_f32, void* stream) {
    int r = dims(height, width, "sgi_kpcn_conv3x3");
    if (r != SGI_OK) return r;
    if (!in || !weights || !out) return fail(SGI_E_INVALID, "sgi_kpcn_conv3x3: NULL buffer");
    if (!aligned16(in) || !aligned16(weights) || !aligned16(out) || (mask && !aligned16(mask)))
        return fail(SGI_E_INVALID, "sgi_kpcn_conv3x3: buffers must be 16-byte aligned");
    const hipError_t e = sgi_kpcn::conv3x3(in, mask, weights, bias, relu ? 1 : 0, height, width, out, out_f32 ? 1 : 0,
                                           (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_kpcn_conv3x3");
}

int sgi_kpcn_conv3x3_fp8(const uint8_t* in, const uint8_t* weights, const float* epilogue, int32_t height,
                         int32_t width, uint8_t* out, void* stream) {
    int r = dims(height, width, "sgi_kpcn_conv3x3_fp8");
    if (r != SGI_OK) return r;
    if (!in || !weights || !epilogue || !out) return fail(SGI_E_INVALID, "sgi_kpcn_conv3x3_fp8: NULL buffer");
    if (!aligned16(in) || !aligned16(weights) || !aligned16(out))
        return fail(SGI_E_INVALID, "sgi_kpcn_conv3x3_fp8: buffers must be 16-byte aligned");
    const hipError_t e = sgi_kpcn::convF8(in, height, width, weights, epilogue, out, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_kpcn_conv3x3_fp8");
}

int sgi_kpcn_bf16_encode(const float* in, int64_t n, uint16_t* out, void* stream) {
    if (n < 0 || n > ((int64_t)1 << 38) || (n > 0 && (!in || !out)))
        return fail(SGI_E_INVALID, "sgi_kpcn_bf16_encode: bad buffer or count");
    if (n == 0) return SGI_OK;
    const hipError_t e = sgi_kpcn::encodeBf16(in, n, out, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_kpcn_bf16_encode");
}

int sgi_kpcn_fp8_encode(const float* in, int64_t n, uint8_t* out, void* stream) {
    if (n < 0 || (n > 0 && (!in || !out))) return fail(SGI_E_INVALID, "sgi_kpcn_fp8_encode: bad buffer or count");
    if (n == 0) return SGI_OK;
    const hipError_t e = sgi_kpcn::encodeF8(in, n, out, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_kpcn_fp8_encode");
}

size_t sgi_kpcn_wgrad_partials(int32_t height, int32_t width) {
    if (height <= 0 || width <= 0) return 0;
    return (size_t)sgi_kpcn::wgradBlocks(height, width) * SGI_KPCN_LAYER_PARAMS;
}

int sgi_kpcn_conv3x3_wgrad(const uint16_t* x, const uint16_t* dy, const uint16_t* mask, int32_t height,
                           int32_t width, float* partials, float* grad, void* stream) {
    int r = dims(height, width, "sgi_kpcn_conv3x3_wgrad");
    if (r != SGI_OK) return r;
    if (!x || !dy || !partials || !grad) return fail(SGI_E_INVALID, "sgi_kpcn_conv3x3_wgrad: NULL buffer");
    if (!aligned16(x) || !aligned16(dy) || (mask && !aligned16(mask)))
        return fail(SGI_E_INVALID, "sgi_kpcn_conv3x3_wgrad: buffers must be 16-byte aligned");
    const hipError_t e = sgi_kpcn::wgrad(x, dy, mask, height, width, partials, grad, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_kpcn_conv3x3_wgrad");
}

int sgi_kpcn_pool(const uint16_t* in, int32_t height, int32_t width, uint16_t* out, void* stream) {
    int r = dims(height, width, "sgi_kpcn_pool");
    if (r != SGI_OK) return r;
    if (!in || !out || !aligned16(in) || !aligned16(out))
        return fail(SGI_E_INVALID, "sgi_kpcn_pool: NULL or unaligned buffer");
    const hipError_t e = sgi_kpcn::pool(in, height, width, out, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_kpcn_pool");
}

int sgi_kpcn_pool_backward(const uint16_t* dout, int32_t height, int32_t width, uint16_t* din, int32_t accumulate,
                           void* stream) {
    int r = dims(height, width, "sgi_kpcn_pool_backward");
    if (r != SGI_OK) return r;
    if (!dout || !din || !aligned16(dout) || !aligned16(din))
        return fail(SGI_E_INVALID, "sgi_kpcn_pool_backward: NULL or unaligned buffer");
    const hipError_t e = sgi_kpcn::poolBackward(dout, height, width, din, accumulate ? 1 : 0, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_kpcn_pool_backward");
}

int sgi_kpcn_upsample_add(const uint16_t* skip, const uint16_t* low, int32_t height, int32_t width, uint16_t* out,
                          void* stream) {
    int r = dims(height, width, "sgi_kpcn_upsample_add");
    if (r != SGI_OK) return r;
    if (!skip || !low || !out || !aligned16(skip) || !aligned16(low) || !aligned16(out))
        return fail(SGI_E_INVALID, "sgi_kpcn_upsample_add: NULL or unaligned buffer");
    const hipError_t e = sgi_kpcn::upAdd(skip, low, height, width, out, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_kpcn_upsample_add");
}

int sgi_kpcn_upsample_backward(const uint16_t* dout, int32_t height, int32_t width, uint16_t* dlow, int32_t accumulate,
                               void* stream) {
    int r = dims(height, width, "sgi_kpcn_upsample_backward");
    if (r != SGI_OK) return r;
    if (!dout || !dlow || !aligned16(dout) || !aligned16(dlow))
        return fail(SGI_E_INVALID, "sgi_kpcn_upsample_backward: NULL or unaligned buffer");
    const hipError_t e =
        sgi_kpcn::upBackward(dout, height, width, dlow, accumulate ? 1 : 0, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_kpcn_upsample_backward");
}

size_t sgi_kpcn_pyramid_floats(int32_t height, int32_t width) {
    if (height <= 0 || width <= 0 || height > 16384 || width > 16384) return 0;
    return sgi_kpcn::pyramidFloats(height, width);
}

int sgi_kpcn_radiance_pyramid(const float* radiance, int32_t height, int32_t width, float* pyramid, void* stream) {
    int r = dims(height, width, "sgi_kpcn_radiance_pyramid");
    if (r != SGI_OK) return r;
    if (!radiance || !pyramid) return fail(SGI_E_INVALID, "sgi_kpcn_radiance_pyramid: NULL buffer");
    const hipError_t e = sgi_kpcn::pyramid(radiance, height, width, pyramid, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_kpcn_radiance_pyramid");
}

int sgi_kpcn_apply_backward(const float* logits, const float* radiance, const float* coarse, const float* dout,
                            const float* derr, int32_t height, int32_t width, uint16_t* dlogits, float* dup,
                            void* stream) {
    int r = dims(height, width, "sgi_kpcn_apply_backward");
    if (r != SGI_OK) return r;
    if (!logits || !radiance || !dout || !dlogits || !aligned16(logits) || !aligned16(dlogits))
        return fail(SGI_E_INVALID, "sgi_kpcn_apply_backward: bad buffer");
    if (coarse && !dup) return fail(SGI_E_INVALID, "sgi_kpcn_apply_backward: a coarser level needs dup");
    const hipError_t e = sgi_kpcn::applyBackward(logits, radiance, coarse, dout, derr, height, width, dlogits, dup,
                                                 (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_kpcn_apply_backward");
}

int sgi_kpcn_upsample_bilinear_backward(const float* dup, int32_t height, int32_t width, float* dcoarse,
                                        void* stream) {
    int r = dims(height, width, "sgi_kpcn_upsample_bilinear_backward");
    if (r != SGI_OK) return r;
    if (!dup || !dcoarse) return fail(SGI_E_INVALID, "sgi_kpcn_upsample_bilinear_backward: NULL buffer");
    const hipError_t e = sgi_kpcn::upBilinearBackward(dup, height, width, dcoarse, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_kpcn_upsample_bilinear_backward");
}

int sgi_kpcn_features(sgi_ctx* ctx, const sgi_gbuffer* g, const float* reservoir, uint16_t* features, float* radiance,
                      void* stream) {
    if (!ctx) return fail(SGI_E_INVALID, "sgi_kpcn_features: NULL context");
    if (!ctx->hasFrame) return fail(SGI_E_STATE, "sgi_kpcn_features: sgi_set_frame first");
    int r = checkGbuffer(ctx, g, "sgi_kpcn_features", true);
    if (r != SGI_OK) return r;
    if (!reservoir || !features || !radiance || !aligned16(reservoir) || !aligned16(features))
        return fail(SGI_E_INVALID, "sgi_kpcn_features: NULL or unaligned buffer");
    const hipError_t e =
        sgi_kpcn::features(sgi_dev::gbufArgs(*g), ctx->frameArgs, reservoir, features, radiance, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_kpcn_features");
}

int sgi_kpcn_apply(const float* logits, const float* radiance, const float* coarse, int32_t height, int32_t width,
                   float* out, float* err, void* stream) {
    int r = dims(height, width, "sgi_kpcn_apply");
    if (r != SGI_OK) return r;
    if (!logits || !radiance || !out || !aligned16(logits)) return fail(SGI_E_INVALID, "sgi_kpcn_apply: bad buffer");
    const hipError_t e = sgi_kpcn::apply(logits, radiance, coarse, height, width, out, err, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_kpcn_apply");
}

int sgi_kpcn_loss(const float* out, const float* logits, const float* target, int32_t height, int32_t width,
                  float err_weight, float scale, float* dout, float* derr, float* loss_sums, void* stream) {
    int r = dims(height, width, "sgi_kpcn_loss");
    if (r != SGI_OK) return r;
    if (!out || !logits || !target || !dout || !derr || !loss_sums)
        return fail(SGI_E_INVALID, "sgi_kpcn_loss: NULL buffer");
    const hipError_t e = sgi_kpcn::loss(out, logits, target, nullptr, height, width, err_weight, scale, dout, derr,
                                        loss_sums, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_kpcn_loss");
}

int sgi_kpcn_loss_ratio(const float* out, const float* logits, const float* target, const float* alternative,
                        int32_t height, int32_t width, float err_weight, float scale, float* dout, float* derr,
                        float* loss_sums, void* stream) {
    int r = dims(height, width, "sgi_kpcn_loss_ratio");
    if (r != SGI_OK) return r;
    if (!out || !logits || !target || !alternative || !dout || !derr || !loss_sums)
        return fail(SGI_E_INVALID, "sgi_kpcn_loss_ratio: NULL buffer");
    const hipError_t e = sgi_kpcn::loss(out, logits, target, alternative, height, width, err_weight, scale, dout,
                                        derr, loss_sums, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_kpcn_loss_ratio");
}

int sgi_kpcn_adam(float* params, const float* grad, float* m, float* v, float lr, float beta1, float beta2, float eps,
                  int32_t step, uint16_t* weights, uint16_t* weights_t, float* biases, void* stream) {
    if (!params || !grad || !m || !v) return fail(SGI_E_INVALID, "sgi_kpcn_adam: NULL buffer");
    if (step < 1) return fail(SGI_E_INVALID, "sgi_kpcn_adam: step counts from 1");
    const hipError_t e = sgi_kpcn::adam(params, grad, m, v, lr, beta1, beta2, eps, step, weights, weights_t, biases,
                                        (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_kpcn_adam");
}

int sgi_kpcn_pack_weights(const float* params, uint16_t* weights, uint16_t* weights_t, float* biases, void* stream) {
    if (!params) return fail(SGI_E_INVALID, "sgi_kpcn_pack_weights: NULL parameters");
    const hipError_t e = sgi_kpcn::packWeights(params, weights, weights_t, biases, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_kpcn_pack_weights");
}

}  // extern "C"

namespace {

// The inference workspace: per level l the activation plane e_l and, below
// level 0, a temporary for c_l / d_l and the pooled input p_l; then the
// features, the radiance pyramid, the reconstruction of every level, the
// tile flags and the list of the tiles that fall back (the level-0 head forms
// the flags from its error logit directly, so no error plane is stored).
struct KpcnWs {
    uint16_t* act[SGI_KPCN_LEVELS][3];  // [l][0] e_l, [l][1] c_l / d_l, [l][2] p_l (l >= 1)
    uint16_t* feat;
    float* rad[SGI_KPCN_LEVELS];  // the radiance pyramid (level 0 = the noisy radiance)
    float* o[SGI_KPCN_LEVELS];    // the reconstruction per level (level 0 = the output)
    uint8_t* flags;
    int32_t* dilatedList;  // [1 + tiles]: count, then tile indices
    int h[SGI_KPCN_LEVELS], w[SGI_KPCN_LEVELS];
};

size_t kpcnLayout(int H, int W, char* base, KpcnWs* ws) {
    size_t off = 0;
    auto take = [&](size_t bytes) -> char* {
        char* p = base ? base + off : nullptr;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    int h = H, w = W;
    for (int l = 0; l < SGI_KPCN_LEVELS; ++l) {
        if (ws) ws->h[l] = h, ws->w[l] = w;
        for (int k = 0; k < 3; ++k) {
            char* p = (k == 0 || l > 0) ? take((size_t)h * w * SGI_KPCN_CHANNELS * 2) : nullptr;
            if (ws) ws->act[l][k] = (uint16_t*)p;
        }
        h = (h + 1) / 2;
        w = (w + 1) / 2;
    }
    const size_t P = (size_t)H * W;
    const size_t tiles = (size_t)((W + 15) / 16) * ((H + 15) / 16);
    char* feat = take(P * 16 * 2);  // the inference features, [h][w][16]
    char* rad0 = take(P * 3 * 4);
    // levels 1..4 of the pyramid contiguous (the layout sgi_kpcn_radiance_pyramid writes)
    char* pyr = take(sgi_kpcn::pyramidFloats(H, W) * 4);
    char* outs[SGI_KPCN_LEVELS];
    h = H, w = W;
    for (int l = 0; l < SGI_KPCN_LEVELS; ++l) {
        outs[l] = take((size_t)h * w * 3 * 4);
        h = (h + 1) / 2;
        w = (w + 1) / 2;
    }
    char* flags = take(tiles);
    char* dilList = take((1 + tiles) * 4);
    if (ws) {
        ws->feat = (uint16_t*)feat;
        ws->rad[0] = (float*)rad0;
        size_t off = 0;
        for (int l = 1; l < SGI_KPCN_LEVELS; ++l) {
            ws->rad[l] = (float*)pyr + off;
            off += (size_t)ws->h[l] * ws->w[l] * 3;
        }
        for (int l = 0; l < SGI_KPCN_LEVELS; ++l) ws->o[l] = (float*)outs[l];
        ws->flags = (uint8_t*)flags;
        ws->dilatedList = (int32_t*)dilList;
    }
    return off;
}

}  // namespace

namespace sgi_launch {

// The body of sgi_denoise_learned (arguments checked by the caller; workspace of
// sgi_kpcn_workspace_bytes(c.H, c.W) bytes): shared with the batched views.
hipError_t denoiseLearned(const sgi_dev::GbufArgs& c, const sgi_dev::GbufArgs& p, const sgi_dev::FrameArgs& fr,
                          const int32_t* visibleIndices, const float* screen, const float* color,
                          const float* reservoir, const float* prevGi, const float* prevHistory,
                          const sgi_kpcn_net& netArgs, float threshold, void* workspace, float* combined,
                          float* giOut, float* reservoirOut, float* historyOut, int32_t* fallbackTiles,
                          const DenoiseBuffers& dws, hipEvent_t screenReady, hipStream_t s) {
    const int H = c.H, W = c.W;
    char* base = (char*)(((uintptr_t)workspace + 255) & ~(uintptr_t)255);
    KpcnWs ws;
    kpcnLayout(H, W, base, &ws);
    const sgi_kpcn_net* net = &netArgs;
    auto Wl = [&](int l) { return net->weights + (size_t)l * SGI_KPCN_LAYER_WEIGHTS; };
    auto Bl = [&](int l) { return net->biases + (size_t)l * SGI_KPCN_CHANNELS; };
    // FP8 inference: layers 1..16 on E4M3 activations and weights, each with its epilogue
    // (k1, k2); layer 0 reads the BF16 features with the BF16 weights and stores E4M3
    const int f8 = net->weights_fp8 != nullptr;
    auto Wt = [&](int l) -> const void* {
        return f8 ? (const void*)(net->weights_fp8 + (size_t)l * SGI_KPCN_LAYER_WEIGHTS) : (const void*)Wl(l);
    };
    auto Kt = [&](int l) { return f8 ? net->epilogue + (size_t)l * 2 * SGI_KPCN_CHANNELS : Bl(l); };
    auto conv = [&](const void* in, int l, int lev, void* out) {
        return f8 ? sgi_kpcn::convF8(in, ws.h[lev], ws.w[lev], Wt(l), Kt(l), out, s)
                  : sgi_kpcn::conv3x3(static_cast<const uint16_t*>(in), nullptr, Wl(l), Bl(l), 1, ws.h[lev], ws.w[lev],
                                      out, 0, s);
    };
    hipError_t e = hipSuccess;
#define KPCN_STEP(call) \
    if (e == hipSuccess) e = (call)
    // encoder: level 0 from the G-buffer features, then pool -> conv -> conv per level
    KPCN_STEP(sgi_kpcn::features16Pyramid(c, fr, reservoir, ws.feat, ws.rad[0], ws.rad[1], s));
    // layer 0 (features -> e0) also writes p1, the 2x2 mean of e0; likewise each level's second
    // convolution writes the next level's input
    KPCN_STEP(sgi_kpcn::conv3x3In16AndPool(ws.feat, ws.h[0], ws.w[0], Wl(0), Kt(0), ws.act[0][0], ws.act[1][2], f8,
                                           s));
    for (int l = 1; l < SGI_KPCN_LEVELS; ++l) {
        KPCN_STEP(conv(ws.act[l][2], 2 * l - 1, l, ws.act[l][1]));
        if (l + 1 < SGI_KPCN_LEVELS) {
            KPCN_STEP(sgi_kpcn::convAndPool(ws.act[l][1], ws.h[l], ws.w[l], Wt(2 * l), Kt(2 * l), ws.act[l][0],
                                            ws.act[l + 1][2], f8, s));
        } else {
            KPCN_STEP(conv(ws.act[l][1], 2 * l, l, ws.act[l][0]));
        }
    }
    // decoder: d_l = conv(e_l + up(d_{l+1})) for l = 3..1 with d_4 = e_4 (d_l lands in act[l][1]);
    // each level's head runs the level's reconstruction step in its epilogue, coarse to fine;
    // the level-0 head reads e0 + up(d1) directly
    constexpr int L4 = SGI_KPCN_LEVELS - 1;
    auto head = [](int l) { return 12 + l; };
    KPCN_STEP(sgi_kpcn::convApply(ws.act[L4][0], ws.h[L4], ws.w[L4], Wt(head(L4)), Kt(head(L4)), ws.rad[L4], nullptr,
                                  ws.o[L4], nullptr, f8, s));
    const void* low = ws.act[L4][0];
    for (int l = L4 - 1, layer = 9; l >= 1; --l, ++layer) {
        KPCN_STEP(sgi_kpcn::convUpAdd(ws.act[l][0], low, ws.h[l], ws.w[l], Wt(layer), Kt(layer), ws.act[l][1], f8, s));
        low = ws.act[l][1];
        KPCN_STEP(sgi_kpcn::convApply(ws.act[l][1], ws.h[l], ws.w[l], Wt(head(l)), Kt(head(l)), ws.rad[l],
                                      ws.o[l + 1], ws.o[l], nullptr, f8, s));
    }
    const bool noFallback = std::isinf(threshold) && threshold > 0.0f;
    // with a threshold the head also forms the per-tile fallback flags
    KPCN_STEP(sgi_kpcn::convUpAddApply(ws.act[0][0], low, H, W, Wt(head(0)), Kt(head(0)), ws.rad[0], ws.o[1], ws.o[0],
                                       nullptr, f8, s, threshold, noFallback ? nullptr : ws.flags,
                                       ws.dilatedList, fallbackTiles));
    if (noFallback) {
        // no threshold: no tile can fall back; the temporal pass reads the network's output
        if (fallbackTiles) KPCN_STEP(sgi_kpcn::clearCounts(H, W, fallbackTiles, s));
        if (screenReady) KPCN_STEP(hipStreamWaitEvent(s, screenReady, 0));
        KPCN_STEP(sgi_launch::denoiseTemporalNet(c, p, fr, screen, color, ws.o[0], prevGi, prevHistory,
                                                 combined, giOut, reservoirOut, historyOut, nullptr, nullptr, s));
    } else {
        KPCN_STEP(sgi_kpcn::dilateFlags(H, W, ws.flags, ws.dilatedList, fallbackTiles, s));
        // the a-trous passes over the flagged tiles and their margin (on a tile list: their cost
        // follows the flagged tiles), then one temporal pass: the a-trous result in the flagged
        // tiles, the network's output in every other tile (which also overwrites the margin's
        // reservoir)
        KPCN_STEP(sgi_launch::denoiseSpatial(c, fr, visibleIndices, reservoir, reservoirOut, dws,
                                             ws.dilatedList, false, s));
        if (screenReady) KPCN_STEP(hipStreamWaitEvent(s, screenReady, 0));
        KPCN_STEP(sgi_launch::denoiseTemporalNet(c, p, fr, screen, color, ws.o[0], prevGi, prevHistory,
                                                 combined, giOut, reservoirOut, historyOut, ws.flags, dws.normals,
                                                 s));
    }
#undef KPCN_STEP
    return e;
}

}  // namespace sgi_launch

extern "C" {

size_t sgi_kpcn_workspace_bytes(int32_t height, int32_t width) {
    if (height <= 0 || width <= 0 || height > 16384 || width > 16384) return 0;
    return kpcnLayout(height, width, nullptr, nullptr) + 256;
}

int sgi_denoise_learned(sgi_ctx* ctx, const sgi_gbuffer* cur, const sgi_gbuffer* prev, const int32_t* visibleIndices,
                        const float* screen, const float* color, const float* reservoir, const float* prevGi,
                        const float* prevHistory, const sgi_kpcn_net* net, float threshold, void* workspace,
                        size_t workspaceBytes, float* combined, float* giOut, float* reservoirOut, float* historyOut,
                        int32_t* fallbackTiles, void* stream) {
    if (!ctx) return fail(SGI_E_INVALID, "sgi_denoise_learned: NULL context");
    if (!ctx->hasFrame) return fail(SGI_E_STATE, "sgi_denoise_learned: sgi_set_frame first");
    int r = checkGbuffer(ctx, cur, "sgi_denoise_learned", true);
    if (r != SGI_OK) return r;
    if (prev && (r = checkGbuffer(ctx, prev, "sgi_denoise_learned(prev)", false)) != SGI_OK) return r;
    if (!visibleIndices || !screen || !color || !reservoir || !combined || !giOut || !reservoirOut || !historyOut)
        return fail(SGI_E_INVALID, "sgi_denoise_learned: NULL buffer");
    if (!net || !net->weights || !net->biases || !aligned16(net->weights))
        return fail(SGI_E_INVALID, "sgi_denoise_learned: network weights missing or unaligned");
    if (net->weights_fp8 && (!net->epilogue || !aligned16(net->weights_fp8)))
        return fail(SGI_E_INVALID, "sgi_denoise_learned: FP8 weights without their epilogue, or unaligned");
    if (!aligned16(reservoir) || !aligned16(reservoirOut))
        return fail(SGI_E_INVALID, "sgi_denoise_learned: reservoirs must be 16-byte aligned");
    const int H = cur->height, W = cur->width;
    if (!workspace || workspaceBytes < sgi_kpcn_workspace_bytes(H, W))
        return fail(SGI_E_INVALID, "sgi_denoise_learned: workspace smaller than sgi_kpcn_workspace_bytes()");
    const hipError_t e = sgi_launch::denoiseLearned(sgi_dev::gbufArgs(*cur), sgi_dev::gbufArgs(prev ? *prev : *cur),
                                                    ctx->frameArgs, visibleIndices, screen, color, reservoir, prevGi,
                                                    prevHistory, *net, threshold, workspace, combined, giOut,
                                                    reservoirOut, historyOut, fallbackTiles, ctx->ws,
                                                    ctx->screenReady, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_denoise_learned");
}

int sgi_render_vpl_cubes(sgi_ctx* ctx, const sgi_triangles* mesh, const sgi_vpl* vpls, const int32_t* updates,
                         int32_t numUpdates, const sgi_atlas_entry* entries, float nearPlane,
                         const sgi_cube_atlases* distance, const sgi_cube_atlases* diffuse, void* stream) {
    const char* who = "sgi_render_vpl_cubes";
    if (!ctx || !mesh || !distance) return fail(SGI_E_INVALID, std::string(who) + ": NULL argument");
    if (numUpdates < 0 || numUpdates > SGI_MAX_TOTAL_VPLS_PER_FRAME)
        return fail(SGI_E_INVALID, std::string(who) + ": numUpdates must be within 0..4480");
    if (mesh->count < 0 || (mesh->count > 0 && !mesh->positions) || mesh->cull < SGI_CULLING_NONE ||
        mesh->cull > SGI_CULLING_CCW)
        return fail(SGI_E_INVALID, std::string(who) + ": bad triangle list");
    if (!(nearPlane > 0.0f) || !std::isfinite(nearPlane))
        return fail(SGI_E_INVALID, std::string(who) + ": nearPlane must be finite and > 0");
    if (numUpdates > 0 && (!vpls || !updates || !entries))
        return fail(SGI_E_INVALID, std::string(who) + ": NULL VPL / update / entry buffer");
    sgi_dev::CubeArgs dz{}, dc{};
    int r = cubeArgs(distance, who, dz);
    if (r != SGI_OK) return r;
    if (diffuse) {
        if (diffuse->format != SGI_TEXEL_RGBA8 || diffuse->resolution != distance->resolution ||
            diffuse->layers_per_atlas != distance->layers_per_atlas || diffuse->num_atlases != distance->num_atlases)
            return fail(SGI_E_INVALID, std::string(who) + ": the diffuse atlases must be RGBA8 of the distance geometry");
        sgi_cube_atlases g = *diffuse;
        g.format = SGI_TEXEL_F32;  // cubeArgs checks the geometry; the texel size is the caller's
        r = cubeArgs(&g, who, dc);
        if (r != SGI_OK) return r;
    }
    if (numUpdates == 0) return SGI_OK;
    if (diffuse && mesh->count > 0 && !mesh->colors)
        return fail(SGI_E_INVALID, std::string(who) + ": diffuse output needs per-triangle colours");
    // per-(light, face) lists of up to count triangles: batches of lights within a 256 MB budget
    const size_t perLight = sgi_launch::cubeListInts(1, mesh->count);
    const int batchMax = (int)std::max<size_t>(1, std::min<size_t>((size_t)numUpdates, (64u << 20) / perLight));
    const size_t need = sgi_launch::cubeListInts(batchMax, mesh->count);
    if (need > ctx->cubeListInts) {
        if (ctx->cubeLists) (void)hipFree(ctx->cubeLists);
        ctx->cubeLists = nullptr;
        ctx->cubeListInts = 0;
        const hipError_t e = hipMalloc((void**)&ctx->cubeLists, need * sizeof(int32_t));
        if (e != hipSuccess) return hipFail(e, "sgi_render_vpl_cubes: list workspace");
        ctx->cubeListInts = need;
    }
    const hipError_t e = sgi_launch::renderVplCubes(mesh->positions, diffuse ? mesh->colors : nullptr, mesh->count,
                                                    mesh->cull, vpls, updates, numUpdates, entries, nearPlane, dz,
                                                    distance->format, dc, diffuse != nullptr, batchMax, ctx->cubeLists,
                                                    (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, who);
}

int sgi_render_csm(sgi_ctx* ctx, const sgi_triangles* mesh, const sgi_csm* csm, float slopeFactor, float* depth,
                   void* stream) {
    const char* who = "sgi_render_csm";
    if (!ctx || !mesh || !csm || !depth) return fail(SGI_E_INVALID, std::string(who) + ": NULL argument");
    if (mesh->count < 0 || (mesh->count > 0 && !mesh->positions) || mesh->cull < SGI_CULLING_NONE ||
        mesh->cull > SGI_CULLING_CCW)
        return fail(SGI_E_INVALID, std::string(who) + ": bad triangle list");
    if (csm->resolution <= 0 || csm->resolution > 16384)
        return fail(SGI_E_INVALID, std::string(who) + ": cascade resolution must be within 1..16384");
    if (!std::isfinite(slopeFactor)) return fail(SGI_E_INVALID, std::string(who) + ": slopeFactor must be finite");
    // the pass reverses each mesh's culling (RenderImmediate_(..., true)): a CCW mesh keeps its clockwise
    // (light-averted) triangles, a CW mesh its counter-clockwise ones
    const int keepSign = mesh->cull == SGI_CULLING_CCW ? -1 : (mesh->cull == SGI_CULLING_CW ? 1 : 0);
    const size_t need = sgi_launch::csmScratchFloats(mesh->count);
    if (need > ctx->csmTrisFloats) {
        if (ctx->csmTris) (void)hipFree(ctx->csmTris);
        ctx->csmTris = nullptr;
        ctx->csmTrisFloats = 0;
        const hipError_t e = hipMalloc((void**)&ctx->csmTris, need * sizeof(float));
        if (e != hipSuccess) return hipFail(e, "sgi_render_csm: triangle workspace");
        ctx->csmTrisFloats = need;
    }
    const hipError_t e = sgi_launch::renderCsm(mesh->positions, mesh->count, keepSign, slopeFactor, csm->resolution,
                                               csm->cascadeProjViews, ctx->csmTris, depth, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, who);
}

int sgi_gather_samples_noalloc(const sgi_frame* frame, const sgi_gbuffer* g, const sgi_vpl* vpls,
                                const int32_t* visibleIndices, const sgi_atlas_entry* shadowIndices,
                                const sgi_cube_atlases* cubes, int32_t maxVisible, int32_t samples, float* color,
                                float* reservoir, void* stream) {
    const char* who = "sgi_gather_samples_noalloc";
    if (!frame || !g || !vpls || !visibleIndices || !shadowIndices || !cubes || !color || !reservoir)
        return fail(SGI_E_INVALID, std::string(who) + ": NULL argument");
    if (maxVisible < 0 || maxVisible > kNoAllocMaxVisible)
        return fail(SGI_E_INVALID, std::string(who) + ": maxVisible must be within 0..4480 (the scratch holds " +
                                       std::to_string(kNoAllocMaxVisible) + " packed VPL records)");
    if (samples < 1 || samples > 4096) return fail(SGI_E_INVALID, std::string(who) + ": samples must be within 1..4096");
    if (g->width <= 0 || g->height <= 0 || g->width > 16384 || g->height > 16384 || !g->depth || !g->normal ||
        !g->albedo || !g->baseReflectivity || !g->roughnessMetallicAmbient)
        return fail(SGI_E_INVALID, std::string(who) + ": bad gbuffer");
    if (!aligned16(reservoir)) return fail(SGI_E_INVALID, std::string(who) + ": reservoir must be 16-byte aligned");
    if (!(frame->framesPerSecond > 0.0f)) return fail(SGI_E_INVALID, std::string(who) + ": framesPerSecond must be > 0");
    sgi_dev::CubeArgs ca;
    int r = cubeArgs(cubes, who, ca);
    if (r != SGI_OK) return r;
    void* prep = nullptr;
    r = noAllocScratch(&prep);
    if (r != SGI_OK) return r;
    const hipError_t e = sgi_launch::gatherEx(sgi_dev::gbufArgs(*g), sgi_dev::frameArgs(*frame), ca, cubes->format,
                                              SGI_GATHER_SAMPLES, samples, -1, 0, 0, vpls, visibleIndices,
                                              shadowIndices, prep, color, reservoir, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, who);
}

}  // extern "C"
