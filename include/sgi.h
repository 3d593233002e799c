/*
 * sgi.h — C ABI of the MI355X deferred VPL global-illumination pass.
 *
 * This library replaces, for one GL frame, the reference renderer's VPL GI
 * pass (StratusGFX Source/Engine/StratusRendererBackend.cpp:1954-2080):
 * global VPL sun-visibility culling, per-tile VPL culling, the per-pixel VPL
 * gather, deferred direct sun lighting with cascaded-shadow PCF, and the
 * 4-pass a-trous / temporal denoiser.  Every call is stream ordered on the
 * HIP stream passed in `stream` (NULL = the default stream); every pointer
 * named "device" is HIP device memory owned by the caller.
 *
 * Texture conventions (restating the reference's GL state):
 *  - 2D planes are row-major [height][width][channels] FP32, row 0 is the
 *    bottom image row (GL window origin); pixel (x, y) is sampled at
 *    uv = ((x + 0.5) / W, (y + 0.5) / H).  Every G-buffer / GI texture in the
 *    reference is CLAMP_TO_EDGE (StratusRendererBackend.cpp:384-589) and is
 *    only ever sampled at texel centres, so a tap is texel (clamp(x), clamp(y)).
 *  - Cube map atlases follow GL_TEXTURE_CUBE_MAP_ARRAY: texel (i, j) of face f
 *    of layer l is at ((l * 6 + f) * res + j) * res + i; faces are
 *    +X, -X, +Y, -Y, +Z, -Z and (s, t) come from the GL major-axis table.
 *    Distance texels hold distance / farPlane (Source/Shaders/shadowVpl.fs:17-35);
 *    diffuse texels hold the nearest surface's base colour as UNORM8 RGB
 *    (shadowVpl.fs:37, the RGB8 colour attachment of
 *    StratusRendererBackend.cpp:2419), read as k / 255.
 *  - Matrices are 4x4 column-major (glm / GL uniform layout).
 *
 * Errors: every entry point returns SGI_OK (0) or a negative SGI_E_* code and
 * records a message readable with sgi_last_error() (thread local).  No C++
 * exception crosses this boundary.  The reference throws std::runtime_error or
 * returns silently (StratusRendererBackend.cpp:1582,1673,1680,1817); the
 * silent "no VPLs" returns are kept as SGI_OK with untouched outputs where the
 * reference leaves its FBOs untouched.
 */
#ifndef SGI_H
#define SGI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Constants shared with the reference shaders (vpl_common.glsl:7-9,
 * pbr.glsl:70-72, StratusGpuCommon.h:15-36). */
#define SGI_MAX_TOTAL_VPLS_BEFORE_CULLING 10000
#define SGI_MAX_TOTAL_SHADOW_ATLASES 14
#define SGI_MAX_TOTAL_SHADOWS_PER_ATLAS 320
#define SGI_MAX_TOTAL_VPLS_PER_FRAME (SGI_MAX_TOTAL_SHADOW_ATLASES * SGI_MAX_TOTAL_SHADOWS_PER_ATLAS)
#define SGI_REFERENCE_MAX_VPLS_PER_TILE 12
#define SGI_MAX_VPLS_PER_TILE_LIMIT 64
#define SGI_NUM_CASCADES 4

enum {
    SGI_OK = 0,
    SGI_E_INVALID = -1,   /* bad argument (NULL, size out of range, ...) */
    SGI_E_HIP = -2,       /* a HIP runtime call failed */
    SGI_E_STATE = -3      /* a required sgi_set_* call is missing */
};

/* GpuVplData (StratusGpuCommon.h:217-232, vpl_common.glsl:48-56), std430, 64 B. */
typedef struct sgi_vpl {
    float position[4];
    float color[4];
    float specularPosition[4];
    float radius;
    float farPlane;
    float intensity;
    float _pad;
} sgi_vpl;

/* GpuAtlasEntry (StratusGpuCommon.h:276-279, pbr.glsl:75-78), 8 B. */
typedef struct sgi_atlas_entry {
    int32_t index; /* atlas (samplerCubeArray) index, < SGI_MAX_TOTAL_SHADOW_ATLASES */
    int32_t layer; /* cube layer inside the atlas */
} sgi_atlas_entry;

/* GpuVplStage1PerTileOutputs (StratusGpuCommon.h:192-195), 32 B. */
typedef struct sgi_tile_stage1 {
    float averageLocalPosition[4];
    float averageLocalNormal[4];
} sgi_tile_stage1;

/* Per-tile VPL list = GpuVplStage2PerTileOutputs (StratusGpuCommon.h:203-209)
 * with the cap K a run-time value: record t is int32[1 + K] at t * (1 + K):
 * { numVisible, indices[K] }.  Entries past numVisible are unspecified. */

struct sgi_gbuffer_packed;

/* G-buffer planes (StratusRendererBackend.h:327-338), FP32.  With `packed`
 * set, every pass reads the normals, albedo, base reflectivity, roughness /
 * metallic / ambient and velocity from the reference's packed attachments
 * instead (the FP32 pointers of those five are then ignored and may be NULL;
 * depth, id and SSAO stay FP32): 22 B per pixel read instead of 60, and the
 * same results bit for bit as the FP32 planes sgi_gbuffer_unpack makes from
 * the same attachments.  packed->width / height must equal width / height. */
typedef struct sgi_gbuffer {
    int32_t width;
    int32_t height;
    const float* depth;                    /* [H][W]    window-space depth in [0,1]   */
    const float* normal;                   /* [H][W][3] normal * 0.5 + 0.5            */
    const float* albedo;                   /* [H][W][4] rgb albedo, a = emissive.r    */
    const float* baseReflectivity;         /* [H][W][2] r = reflectance, g = emissive.g */
    const float* roughnessMetallicAmbient; /* [H][W][3] roughness, metallic, emissive.b */
    const float* velocity;                 /* [H][W][2] uv motion (cur - prev); NULL = 0 */
    const float* id;                       /* [H][W]    mesh id; NULL = 0               */
    const float* ssao;                     /* [H][W]    SSAO; NULL = 1.0                */
    const struct sgi_gbuffer_packed* packed; /* NULL: the FP32 planes above            */
} sgi_gbuffer;

/* Texel formats of the cube atlases.  Distance cubes are F32, F16 (IEEE
 * binary16) or UNORM16.  UNORM16 is the reference's own VPL cube format: the
 * depth attachment of CreateShadowMap3DCache_ (DEPTH, BITS_16,
 * StratusRendererBackend.cpp:296,2412) is GL_DEPTH_COMPONENT16
 * (StratusTexture.cpp:542), whose code k reads as k / 65535 (correctly rounded
 * to FP32), so an engine's atlases can be passed in unchanged.  The diffuse
 * (colour) cubes are RGBA8, 4 bytes per texel (r, g, b UNORM8 at bytes 0..2,
 * byte 3 unused), the 4-byte texel GL keeps for an RGB8 attachment. */
enum { SGI_TEXEL_F32 = 0, SGI_TEXEL_F16 = 1, SGI_TEXEL_RGBA8 = 2, SGI_TEXEL_UNORM16 = 3 };

/* The G-buffer attachments as the reference allocates them
 * (StratusRendererBackend.cpp:384-424), row-major [H][W] like the FP32 planes:
 * 8-bit UNORM channels read as k / 255, RGB8 attachments in 4-byte texels
 * (r, g, b in bytes 0..2, byte 3 unused; the texel GL keeps for RGB8), RG8 in
 * 2-byte texels, velocity as two IEEE binary16 values.  Depth and ids are read
 * as FP32 values and need no conversion: the FP32 sgi_gbuffer takes them as
 * they are. */
typedef struct sgi_gbuffer_packed {
    int32_t width;
    int32_t height;
    const uint32_t* normals;                  /* RGB8:  n * 0.5 + 0.5            */
    const uint32_t* albedo;                   /* RGBA8: rgb albedo, a = emissive.r */
    const uint16_t* baseReflectivity;         /* RG8:   reflectance, emissive.g   */
    const uint32_t* roughnessMetallicAmbient; /* RGB8:  roughness, metallic, emissive.b */
    const uint16_t* velocity;                 /* RG16F; NULL = no velocity plane */
} sgi_gbuffer_packed;

/* VPL cube-map atlases, distance or diffuse (StratusRendererBackend.cpp:2397-2449). */
typedef struct sgi_cube_atlases {
    int32_t resolution;       /* face edge in texels */
    int32_t format;           /* distance: SGI_TEXEL_F32, _F16 or _UNORM16; diffuse: SGI_TEXEL_RGBA8 */
    int32_t layers_per_atlas; /* cube layers in each atlas */
    int32_t num_atlases;      /* <= SGI_MAX_TOTAL_SHADOW_ATLASES */
    const void* atlas[SGI_MAX_TOTAL_SHADOW_ATLASES]; /* device [layers][6][res][res] */
} sgi_cube_atlases;

/* Cascaded sun shadow map (StratusRendererBackend.cpp:341-356, 2452-2474). */
typedef struct sgi_csm {
    int32_t resolution;                      /* square cascade edge in texels */
    const float* depth;                      /* device [4][res][res], depth in [0,1] */
    float cascadeProjViews[SGI_NUM_CASCADES][16];
    float cascadePlanes[SGI_NUM_CASCADES - 1][4];
    float znear;      /* infiniteLightZnear */
    float zfar;       /* infiniteLightZfar  */
    float depthBias;  /* infiniteLightDepthBias */
} sgi_csm;

/* Per-frame uniforms (RendererFrame, StratusRendererBackend.h:297-323, and the
 * renderer settings StratusRendererBackend.h:186-294). */
typedef struct sgi_frame {
    float invProjectionView[16];
    float prevInvProjectionView[16];
    float viewPosition[3];
    float infiniteLightDirection[3]; /* unit vector pointing toward the sun */
    float infiniteLightColor[3];     /* sun colour * intensity (StratusLight.h:62) */
    float time;                      /* RNG seed z (vpl_pbr_gi.fs:132) */
    float framesPerSecond;           /* history cap (vpl_pbr_gi_denoise.fs:422) */
    int32_t frameCount;
    float minRoughness;              /* 0.08 */
    int32_t usePerceptualRoughness;  /* 1 */
    float minGiOcclusionFactor;      /* 0.95 */
    float fogDensity;                /* 0 */
    float fogColor[3];
    float emissionStrength;          /* 0 */
} sgi_frame;

enum { SGI_PRECISION_FP32 = 0, SGI_PRECISION_FP16 = 1 };

typedef struct sgi_config {
    int32_t width;
    int32_t height;
    int32_t tile_size;          /* tile edge in pixels: 8, 16 or 32 */
    int32_t max_vpls_per_tile;  /* K: 1 .. SGI_MAX_VPLS_PER_TILE_LIMIT */
} sgi_config;

typedef struct sgi_ctx sgi_ctx;

/* ---- lifetime ------------------------------------------------------------ */

/* Creates a context for one viewport size (RendererBackend::UpdateWindowDimensions_,
 * StratusRendererBackend.cpp:436-589): validates the config and allocates the
 * denoiser's ping-pong workspace on the current device. */
int sgi_create(const sgi_config* config, sgi_ctx** out);
void sgi_destroy(sgi_ctx* ctx);
const char* sgi_last_error(void);
/* Number of tiles along x / y for the context's config (ceil division). */
int sgi_tile_counts(const sgi_ctx* ctx, int32_t* tiles_x, int32_t* tiles_y);

/* Per-frame state (the uniform uploads of InitCoreCSMData_ / InitLights_,
 * StratusRendererBackend.cpp:2452-2604).  The csm depth pointer is kept, not copied. */
int sgi_set_frame(sgi_ctx* ctx, const sgi_frame* frame);
int sgi_set_csm(sgi_ctx* ctx, const sgi_csm* csm);

/* ---- kernels ------------------------------------------------------------- */

/* Samples the reference's packed G-buffer attachments into the FP32 planes
 * that every pass reads (the texel-centre, NEAREST-equivalent reads the GLSL
 * passes make of these attachments: UNORM8 -> k / 255, binary16 -> FP32, both
 * exact in FP32).  Outputs: device normal [H][W][3], albedo [H][W][4],
 * baseReflectivity [H][W][2], roughnessMetallicAmbient [H][W][3], velocity
 * [H][W][2] (may be NULL when packed->velocity is NULL).  The packed size must
 * be the context's viewport. */
int sgi_gbuffer_unpack(sgi_ctx* ctx, const sgi_gbuffer_packed* packed, float* normal, float* albedo,
                       float* baseReflectivity, float* roughnessMetallicAmbient, float* velocity, void* stream);

/* The geometry pass's attachment writes: its fragment outputs gNormal,
 * gAlbedo, gReflectivityEmissive, gRoughnessMetallicAmbient and
 * gVelocityBuffer (pbr_geometry_pass.fs:127-137), given as the FP32 planes of
 * `planes`, stored into the attachments' formats
 * (StratusRendererBackend.cpp:384-424) as GL converts a fragment output: a
 * UNORM8 channel clamps to [0, 1] and stores the code floor(v * 255 + 0.5)
 * (NaN stores 0), velocity rounds to the nearest binary16 (ties to even);
 * RGB8 texels keep byte 3 zero.  Writes the attachments `out` names (its
 * pointers are the destinations): normals, albedo, baseReflectivity,
 * roughnessMetallicAmbient, and velocity when both planes->velocity and
 * out->velocity are set.  planes->depth, id, ssao and packed are not read.
 * Sizes: the context's viewport.  sgi_gbuffer_unpack(sgi_gbuffer_pack(x))
 * is x rounded to the attachments' precision; sgi_gbuffer_pack of unpacked
 * attachments gives the same codes back. */
int sgi_gbuffer_pack(sgi_ctx* ctx, const sgi_gbuffer* planes, const sgi_gbuffer_packed* out, void* stream);

/* viscull_vpls.cs:47-120 driven by PerformVirtualPointLightCullingStage1_
 * (StratusRendererBackend.cpp:1578-1661).  For i < n, VPL i is sun-visible when
 * 1 - calculateInfiniteShadowValue(position, blends, sunDir, false) < 1.
 * Visible records are compacted into vplUpdatedData in ascending input order
 * and their input indices into visibleIndices[1..]; visibleIndices[0] receives
 * min(count, SGI_MAX_TOTAL_VPLS_PER_FRAME).  (The reference compacts with
 * atomics, so its order is nondeterministic; ascending order is one of the
 * orders it can produce.)  Needs sgi_set_frame and sgi_set_csm.
 * vplData: device [n]; vplUpdatedData: device [min(n, 4480)] (both 16-byte
 * aligned); visibleIndices: device [1 + min(n, 4480)].  0 <= n <= 10000. */
int sgi_cull_global(sgi_ctx* ctx, const sgi_vpl* vplData, int32_t n,
                    sgi_vpl* vplUpdatedData, int32_t* visibleIndices, void* stream);

/* PerformVirtualPointLightCullingStage2_'s host loop (StratusRendererBackend.cpp:
 * 1675-1701), which reads the visible list back and uploads one atlas entry per
 * visible slot, done on the device: out[s] = shadowIndices[visibleIndices[1 + s]]
 * for s < min(visibleIndices[0], maxCount).  shadowIndices: device [n] by input
 * index; out: device [maxCount]. */
int sgi_visible_atlas_entries(sgi_ctx* ctx, const sgi_atlas_entry* shadowIndices,
                              const int32_t* visibleIndices, int32_t maxCount,
                              sgi_atlas_entry* out, void* stream);

/* vpl_light_color.cs:41-76, dispatched by PerformVirtualPointLightCullingStage2_
 * (StratusRendererBackend.cpp:1703-1731) over the compacted visible VPLs.  For
 * each slot s < min(visibleIndices[0], maxCount), with e = visibleEntries[s]
 * (sgi_visible_atlas_entries' output) and d = -infiniteLightDirection, the
 * NEAREST texels of the diffuse and distance cubes of atlas e.index, layer
 * e.layer, in direction d give
 *   vplUpdatedData[s].color = (rgb * infiniteLightColor / 1 * intensity * colorMultiplier, 1)
 *   vplUpdatedData[s].specularPosition = (position - 1.1 * (radius * distance) * infiniteLightDirection, 1)
 * with every product and difference rounded on its own, in the reference's
 * order; the other fields are left as they are.  The reference's multiplier is
 * 50000 (vpl_light_color.cs:44).  diffuse: SGI_TEXEL_RGBA8 atlases; distance:
 * F32, F16 or UNORM16; both of one atlas geometry (the reference keeps them as the two
 * attachments of one framebuffer).  Needs sgi_set_frame.  vplUpdatedData:
 * device [maxCount]; visibleEntries: device [maxCount]; 0 <= maxCount <= 4480. */
int sgi_color_vpls(sgi_ctx* ctx, const int32_t* visibleIndices, const sgi_atlas_entry* visibleEntries,
                   const sgi_cube_atlases* diffuse, const sgi_cube_atlases* distance, float colorMultiplier,
                   int32_t maxCount, sgi_vpl* vplUpdatedData, void* stream);

/* Face culling of a triangle list, RenderFaceCulling (StratusRenderComponents.h:
 * 20-24) as SetCullState applies it (StratusRendererBackend.cpp:901-918):
 * NONE draws both sides; CW / CCW name the front faces' window-space winding
 * and cull the back faces.  Meshes default to CCW (StratusRenderComponents.h:142). */
enum { SGI_CULLING_NONE = 0, SGI_CULLING_CW = 1, SGI_CULLING_CCW = 2 };

/* World-space triangles of the static meshes a shadow pass draws (the
 * reference draws staticPbrMeshes with their model matrices applied in
 * shadow.vs:31 and csm.vs:40, StratusRendererBackend.cpp:1529-1536, 1137). */
typedef struct sgi_triangles {
    int32_t count;
    const float* positions;  /* device [count][3][3]: world-space vertices, wound as the mesh */
    const uint32_t* colors;  /* device [count] RGBA8 base colour (r in byte 0); needed for diffuse output */
    int32_t cull;            /* SGI_CULLING_* of the meshes */
} sgi_triangles;

/* PointLight::lightNearPlane (StratusLight.h:291), the near plane of every VPL cube face. */
#define SGI_REFERENCE_VPL_NEAR_PLANE 0.1f

/* The VPL shadow pre-pass (the cube update loop of
 * StratusRendererBackend.cpp:1426-1575 with shadowVpl.fs:17-37 and
 * shadow.vs:22-33): for each s < numUpdates, light v = vpls[updates[s]] is
 * rendered into cube entries[updates[s]] (layer * 6 + face of that atlas) of
 * the distance atlases and, when diffuse is not NULL, of the diffuse atlases.
 * Per face f, the view is lookAt(pos, pos + axis_f, up_f) (:873-886) with
 * perspective(90 deg, 1, nearPlane, v.farPlane) (:1450); the layer is
 * cleared (distance 1, colour 0) and every triangle drawn (culled by
 * mesh->cull, as the pass draws with the meshes' own modes, :1536) with
 * fragment value saturate(|P - pos| / farPlane) and its base colour, depth test LESS: per
 * texel the smallest stored value (the distance format's F32 value, binary16
 * or UNORM16 code) over the triangles the texel's pixel-centre ray meets with
 * view depth in [nearPlane, farPlane], the first such triangle in list order
 * on ties giving the colour (alpha 255).  The caller plays the reference's
 * light-update queue: updates lists the lights popped this frame (the
 * reference pops at most maxShadowUpdatesPerFrame = 5, StratusRendererBackend.h:
 * 378) or every light for an initial fill.  The reference also draws its
 * skybox into texels nothing covers (:1555-1570); there is no skybox here, so
 * they keep the clear.  An entry outside the atlases is skipped.
 * distance: F32, F16 or UNORM16 (the reference's GL_DEPTH_COMPONENT16);
 * diffuse: RGBA8 atlases of the same geometry, or NULL.  nearPlane:
 * SGI_REFERENCE_VPL_NEAR_PLANE in the reference.  The context keeps a list
 * workspace (about 24 B per triangle per light of a batch, batches within
 * 256 MB), allocated on first use. */
int sgi_render_vpl_cubes(sgi_ctx* ctx, const sgi_triangles* mesh, const sgi_vpl* vpls, const int32_t* updates,
                         int32_t numUpdates, const sgi_atlas_entry* entries, float nearPlane,
                         const sgi_cube_atlases* distance, const sgi_cube_atlases* diffuse, void* stream);

/* glPolygonOffset's slope factor of the cascade render (StratusRendererBackend.cpp:1097). */
#define SGI_REFERENCE_CSM_SLOPE_FACTOR 2.0f

/* The sun cascades' depth render, RenderCSMDepth_ (StratusRendererBackend.cpp:
 * 1081-1158) with csm.vs:29-41: layer c of depth (device [4][res][res] FP32,
 * res = csm->resolution) receives every triangle transformed by
 * csm->cascadeProjViews[c] (the cascade's projectionViewRender): the window
 * depth z * 0.5 + 0.5 at each texel centre, plus glPolygonOffset(slopeFactor,
 * 0)'s term slopeFactor * max(|dz_w/dx_w|, |dz_w/dy_w|) (GL's permitted form of
 * the depth slope), clamped to [0, 1] (GL_DEPTH_CLAMP: no near / far clipping),
 * depth test LESS against the clear 1.  Each mesh's culling is reversed as the
 * pass draws it (:1029-1042, 1136-1137): CCW meshes keep their clockwise
 * (light-averted) sides.  Coverage: the NDC edge functions of the texel centre
 * all of one sign.  The shader's alpha test and the dynamic meshes do not
 * apply (static, opaque triangles).  slopeFactor:
 * SGI_REFERENCE_CSM_SLOPE_FACTOR in the reference.  The context keeps a
 * workspace of 160 B per triangle, allocated on first use. */
int sgi_render_csm(sgi_ctx* ctx, const sgi_triangles* mesh, const sgi_csm* csm, float slopeFactor, float* depth,
                   void* stream);

/* vpl_tiled_deferred_culling_stage1.cs:50-90 + _stage2.cs:61-129 (the dispatches
 * commented out at StratusRendererBackend.cpp:1733-1783).
 * Stage 1: per tile, the mean world position and the normalised mean decoded
 * normal of its in-viewport pixels, summed sequentially in row-major local order.
 * Stage 2: per tile, over slots s = 0 .. count-1 of the visible list, the VPLs
 * vpls[s] with dot(n, L - P) >= 0 and |L - P| / radius <= 1, kept as the K
 * smallest ratios, ties in slot order (the reference's strict-< insertion sort).
 * count = visibleIndices[0].  List entries are slots s when remap == 0, or
 * visibleIndices[1 + s] (the input VPL index, as the reference stage 2 stores)
 * when remap != 0.  stage1 (device [tiles]) may be NULL.
 * tiles: device int32 [tiles_x * tiles_y][1 + K]. */
int sgi_cull_tiles(sgi_ctx* ctx, const sgi_gbuffer* gbuffer, const sgi_vpl* vpls,
                   const int32_t* visibleIndices, int32_t remap,
                   sgi_tile_stage1* stage1, int32_t* tiles, void* stream);

/* vpl_pbr_gi.fs:81-216 (ComputeVirtualPointLightGlobalIllumination_,
 * StratusRendererBackend.cpp:1816-1877).  vpls = vplUpdatedData (compacted),
 * shadowIndices[slot] = that VPL's cube atlas entry, count = visibleIndices[0].
 * tiles == NULL: the reference's live estimator (hashed-random VPL slots from the
 * whole visible list, <= 5 rejections).  tiles != NULL: every VPL of the
 * pixel's tile list is evaluated once, with slot entries (remap == 0).
 * color: device [H][W][3] = albedo + 1e-5; reservoir: device [H][W][4] =
 * (min(sum C, 65504), max(sum w, 1)).  precision: SGI_PRECISION_FP32 (the
 * restatement) or SGI_PRECISION_FP16: the same accept / shadow decisions and
 * the same sum w (both FP32), the BRDF and attenuation of two VPLs at a time in
 * packed FP16, accumulated into FP32 colour; it assumes VPL radii <= 16384. */
int sgi_gather(sgi_ctx* ctx, const sgi_gbuffer* gbuffer, const sgi_vpl* vpls,
               const int32_t* visibleIndices, const sgi_atlas_entry* shadowIndices,
               const sgi_cube_atlases* cubes, const int32_t* tiles, int32_t precision,
               float* color, float* reservoir, void* stream);

/* Gather modes of sgi_gather_ex. */
enum {
    SGI_GATHER_REFERENCE = 0, /* the reference's estimator (as sgi_gather with tiles == NULL) */
    SGI_GATHER_TILES = 1,     /* every VPL of the pixel's tile list (as sgi_gather with tiles) */
    SGI_GATHER_ALL = 2,       /* every visible VPL: the converged estimate of the pass */
    SGI_GATHER_SAMPLES = 3    /* `samples` hashed-random slots over all visible VPLs per pixel */
};

typedef struct sgi_gather_opts {
    int32_t mode;             /* SGI_GATHER_* */
    int32_t samples;          /* SGI_GATHER_SAMPLES: draws per pixel (1..4096) */
    int32_t precision;        /* SGI_PRECISION_* of the shading (FP16: SGI_GATHER_SAMPLES with
                                 rejection_limit < 0 only, draws shaded in pairs) */
    int32_t accumulate_fp64;  /* SGI_GATHER_ALL: sum C and w in FP64 */
    int32_t rejection_limit;  /* SGI_GATHER_SAMPLES, >= 0: a draw failing the side / radius
                                 test is drawn again, and after this many rejections draws are
                                 taken unconditionally (the reference's rule is 5); negative:
                                 exactly `samples` draws, failing ones skipped (no light, no
                                 weight), so the estimate converges to SGI_GATHER_ALL's */
} sgi_gather_opts;

/* The gather's further modes, for measurement and for the learned denoiser's
 * data (no counterpart in the reference, which only has the estimator of
 * vpl_pbr_gi.fs:145-200).  SGI_GATHER_ALL evaluates every visible VPL with the
 * per-VPL body of vpl_pbr_gi.fs:153-199, skipping the ones the side / radius
 * test rejects; SGI_GATHER_SAMPLES draws slot int(count * rand) with the
 * reference's hash and <= 5 rejections.  Outputs as sgi_gather. */
int sgi_gather_ex(sgi_ctx* ctx, const sgi_gbuffer* gbuffer, const sgi_vpl* vpls,
                  const int32_t* visibleIndices, const sgi_atlas_entry* shadowIndices,
                  const sgi_cube_atlases* cubes, const int32_t* tiles,
                  const sgi_gather_opts* opts, float* color, float* reservoir, void* stream);

/* No counterpart in the reference: sgi_gather_ex's fixed-count sampler
 * (SGI_GATHER_SAMPLES, rejection_limit < 0, FP32 shading) for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It needs no context and calls no
 * allocating routine: the per-frame frame block is passed in, and the packed
 * visible-VPL records its kernels use live in a file-scope device array of
 * 4480 x 72 B (315 KB per GPU, held until the process ends).  Every call on a
 * GPU shares that GPU's array, so calls that overlap on one GPU (other streams,
 * other threads) race; each call rewrites the records it reads before reading
 * them.  maxVisible: the capacity of the caller's visible list (the size of
 * vpls and of visibleIndices minus one); the scratch holds 4480 records, so a
 * capacity above 4480 is refused, and the kernels read at most
 * min(visibleIndices[0], 4480) records as sgi_gather_ex does.  Other arguments
 * and the outputs are those of sgi_gather_ex, bit for bit the same results. */
int sgi_gather_samples_noalloc(const sgi_frame* frame, const sgi_gbuffer* gbuffer, const sgi_vpl* vpls,
                               const int32_t* visibleIndices, const sgi_atlas_entry* shadowIndices,
                               const sgi_cube_atlases* cubes, int32_t maxVisible, int32_t samples, float* color,
                               float* reservoir, void* stream);

/* pbr.fs:82-145 with INFINITE_LIGHTING_ENABLED (StratusRendererBackend.cpp:2010-2038),
 * zero point lights: sun lighting (BRDF_Burley) times the 3x3 PCF cascade
 * shadow, plus emissive * emissionStrength, bounded to 65504.
 * lit: device [H][W][3].  Needs sgi_set_frame and sgi_set_csm. */
int sgi_direct_light(sgi_ctx* ctx, const sgi_gbuffer* gbuffer, float* lit, void* stream);

/* vpl_pbr_gi_denoise.fs:142-435 run as the 4-pass loop of
 * StratusRendererBackend.cpp:1912-1951: pass 0 = reservoir merge (centre only)
 * and normalisation, passes 1-3 = 5x5 edge-stopping taps at stride 2^pass,
 * last pass = temporal blend with history and the combine.
 * screen = direct light [H][W][3]; color / reservoir = sgi_gather outputs;
 * prev = previous frame's G-buffer (normal, id, depth are read; may equal cur);
 * prevGi [H][W][3] and prevHistory [H][W] are the previous frame's giOut and
 * historyOut (NULL = zero history).  Outputs: combined [H][W][3] (the lit
 * buffer), giOut [H][W][3], reservoirOut [H][W][4], historyOut [H][W].
 * numVisible = visibleIndices[0] (pass 0's probabilistic weight). */
int sgi_denoise(sgi_ctx* ctx, const sgi_gbuffer* cur, const sgi_gbuffer* prev,
                const int32_t* visibleIndices, const float* screen,
                const float* color, const float* reservoir,
                const float* prevGi, const float* prevHistory,
                float* combined, float* giOut, float* reservoirOut, float* historyOut,
                void* stream);

/* No reference counterpart (the reference renders the direct light and the GI
 * passes in sequence, StratusRendererBackend.cpp:1982-2045).  event: a
 * hipEvent_t that the caller records on another stream after the pass that
 * writes the direct light (screen), or NULL (the default) for none.  With an
 * event set, sgi_denoise, sgi_denoise_learned and sgi_render_frame wait for
 * it (hipStreamWaitEvent on their own stream) just before their temporal pass,
 * the only pass that reads screen, so the direct light may run beside culling,
 * the gather and the spatial passes instead of being joined before the
 * denoiser.  The context keeps the handle: the caller records the event before
 * each such call and keeps it alive while it is set.  Under stream capture the
 * wait becomes a graph edge from the event's record. */
int sgi_set_screen_event(sgi_ctx* ctx, void* event);

/* taa.fs:12-100 (PerformTaaPostFx_, StratusRendererBackend.cpp:2326-2352), the
 * first consumer of the lit buffer: per pixel, with uv the texel centre and
 * vel the velocity texel, the previous frame sampled bilinearly (LINEAR,
 * CLAMP_TO_EDGE) at uv - vel is clamped, in the tonemapped space
 * c / (c + 1), to the range of the current texel and its 4 neighbours
 * (edges clamped); current and clamped history are mapped back with
 * c / (1 - c) and mixed: out = cur * (1 - 0.9) + hist * 0.9.  screen, prevScreen,
 * out: device [H][W][3] FP32 of the viewport (out aliases neither input;
 * the reference's history is its RGB8 previousFrameBuffer, :451-453, whose
 * values k / 255 the caller passes as FP32); velocity: the G-buffer's [H][W][2]
 * plane, NULL = 0.  The next frame's prevScreen is sgi_taa_history(out). */
int sgi_taa(sgi_ctx* ctx, const float* screen, const float* prevScreen, const float* velocity, float* out,
            void* stream);

/* The history copy at the end of PerformTaaPostFx_ (StratusRendererBackend.cpp:
 * 2342-2349): the TAA output, held by the reference in an RGBA16F target
 * (:708-711), blitted NEAREST into the RGB8 UNORM previousFrameBuffer
 * (:449-453).  Per channel: binary16 rounding (to nearest even), clamp to
 * [0, 1], the UNORM8 store floor(v * 255 + 0.5), read back as k / 255 into
 * prevScreen (FP32, what sgi_taa takes).  HDR values above 1 are clamped, as
 * the reference's history clamps them.  taaOut, prevScreen: device [H][W][3]
 * of the viewport; they may alias (the kernel is elementwise). */
int sgi_taa_history(sgi_ctx* ctx, const float* taaOut, float* prevScreen, void* stream);

/* gammaTonemap.fs:125-142 (PerformGammaTonemapPostFx_, StratusRendererBackend.cpp:
 * 2356-2366), the consumer of the lit buffer: per channel the ACES film fit
 * saturate(c (2.51 c + 0.03) / (c (2.43 c + 0.59) + 0.14)), then pow(., 1 / gamma),
 * stored to the RGBA8 UNORM target as floor(saturate(v) * 255 + 0.5), alpha 255.
 * lit: device [H][W][3] FP32 (the viewport); out: device [H][W] RGBA8 texels
 * (r in byte 0).  gamma: 2.2 in the reference.  (The reference's TAA runs
 * before it and FXAA after it, both built here; its bloom and atmospheric
 * passes between TAA and the tonemap are not.) */
int sgi_tonemap(sgi_ctx* ctx, const float* lit, float gamma, uint32_t* out, void* stream);

/* The uniforms of fxaa_smoothing.fs (:72-101), which the reference never sets
 * (their defaults 0.0625, 0.125 and 1.0 apply; pass NULL for them). */
typedef struct sgi_fxaa_params {
    float contrastThreshold;
    float relativeThreshold;
    float subpixelBlending;
} sgi_fxaa_params;

/* PerformFxaaPostFx_ (StratusRendererBackend.cpp:2293-2321), the last pass of
 * the chain: fxaa_luminance.fs (the tonemapped colour with the luminance
 * 0.2126 r + 0.7152 g + 0.0722 b of its saturated rgb, aa_common.glsl:7-9,
 * stored as an RGBA8 target) then fxaa_smoothing.fs over that target (LINEAR,
 * CLAMP_TO_EDGE): edge detection on the plus-shaped neighbourhood's contrast,
 * the sub-pixel blend factor, the walk along the edge (steps 1, 1.5, 2 x 7, 4,
 * then a guess jump of 8 texels) and one bilinear tap offset across the edge,
 * stored RGBA8 with alpha 255.  Both passes run in one kernel: each tap
 * recomputes the luminance target's alpha from the tonemapped codes.
 * tonemapped, out: device [H][W] RGBA8 texels of the viewport (sgi_tonemap's
 * output; r in byte 0), not aliased.  params: NULL = the reference's values. */
int sgi_fxaa(sgi_ctx* ctx, const uint32_t* tonemapped, const sgi_fxaa_params* params, uint32_t* out, void* stream);

/* ---- one frame ------------------------------------------------------------ */

/* Caller-owned buffers of one GI frame, in the order RenderScene uses them. */
typedef struct sgi_frame_io {
    const sgi_gbuffer* gbuffer;
    const sgi_gbuffer* prevGbuffer;        /* NULL = gbuffer */
    const sgi_vpl* vplData;                /* [numVpls] frustum-culled VPLs */
    int32_t numVpls;
    const sgi_atlas_entry* shadowIndices;  /* [numVpls] by input index */
    const sgi_cube_atlases* cubes;
    int32_t useTiles;                      /* 0 = reference random gather, 1 = tile lists */
    int32_t precision;                     /* SGI_PRECISION_* of the gather */
    /* scratch / outputs */
    sgi_vpl* vplUpdatedData;               /* [min(numVpls,4480)] */
    int32_t* visibleIndices;               /* [1 + min(numVpls,4480)] */
    sgi_atlas_entry* visibleShadowIndices; /* [min(numVpls,4480)] */
    int32_t* tiles;                        /* [tiles][1 + K] (useTiles) */
    float* direct;                         /* [H][W][3] */
    float* color;                          /* [H][W][3] */
    float* reservoir;                      /* [H][W][4] */
    const float* prevGi;                   /* history in (NULL = none) */
    const float* prevHistory;
    float* combined;                       /* [H][W][3] lit output */
    float* giOut;                          /* [H][W][3] history out */
    float* reservoirOut;                   /* [H][W][4] */
    float* historyOut;                     /* [H][W]    */
    /* VPL colouring (vpl_light_color.cs): NULL = the records arrive coloured */
    const sgi_cube_atlases* diffuseCubes;  /* RGBA8, the geometry of cubes */
    float vplColorMultiplier;              /* 50000 in the reference */
} sgi_frame_io;

/* RenderScene's GI steps (StratusRendererBackend.cpp:1982, 2010-2045) without
 * host round trips: cull_global -> gather the visible atlas entries ->
 * (color_vpls, when diffuseCubes is set) -> direct_light -> (cull_tiles) ->
 * gather -> denoise. */
int sgi_render_frame(sgi_ctx* ctx, const sgi_frame_io* io, void* stream);

/* ---- batched views ---------------------------------------------------------
 *
 * Many viewpoints of one scene (shared VPLs and cubes), each with its own
 * viewport size, G-buffer, uniforms and cascades, rendered as sgi_render_frame
 * with tile lists would render each view, but with one launch per stage over
 * the work of every view (the stages' blocks are numbered over all views'
 * tiles / pixel blocks), and optionally replayed from a captured graph.  The
 * reference renders one view per frame (RenderScene); this has no counterpart
 * there and serves callers that render many views per frame (probes,
 * reflections, multi-camera capture). */
typedef struct sgi_views sgi_views;

typedef struct sgi_views_config {
    int32_t max_views;          /* 1..64 */
    int64_t max_pixels;         /* total pixels of a batch (workspace size) */
    int32_t tile_size;          /* 8, 16 or 32 */
    int32_t max_vpls_per_tile;  /* 1..64 */
} sgi_views_config;

typedef struct sgi_view_io {
    sgi_gbuffer gbuffer;        /* the view's device planes and viewport size */
    sgi_frame frame;            /* the view's uniforms (as sgi_set_frame) */
    sgi_csm csm;                /* the view's cascades (as sgi_set_csm) */
    int32_t* tiles;             /* device [tiles][1 + K] at the batch's tile size */
    float* direct;              /* device [H][W][3] */
    float* color;               /* device [H][W][3] */
    float* reservoir;           /* device [H][W][4], 16-byte aligned */
    const float* prevGi;        /* history in (NULL = none) */
    const float* prevHistory;
    float* combined;            /* device [H][W][3] */
    float* giOut;               /* device [H][W][3] */
    float* reservoirOut;        /* device [H][W][4], 16-byte aligned */
    float* historyOut;          /* device [H][W] */
} sgi_view_io;

int sgi_views_create(const sgi_views_config* config, sgi_views** out);
void sgi_views_destroy(sgi_views* views);

/* Describes the batch: numViews views, the shared VPLs (input order, as
 * sgi_cull_global), their atlas entries by input index and the cube atlases.
 * The per-view descriptors are copied to the device on `stream`. */
int sgi_views_set(sgi_views* views, int32_t numViews, const sgi_view_io* io, const sgi_vpl* vpls,
                  int32_t numVpls, const sgi_atlas_entry* shadowIndices, const sgi_cube_atlases* cubes,
                  void* stream);

/* One frame of every view: VPL culling, atlas entries, direct light, tile
 * culling, tile-list gather (FP32) and the a-trous / temporal denoiser, each
 * stage one launch over all views. */
int sgi_views_run(sgi_views* views, void* stream);

/* Captures sgi_views_run's launches on `stream` (not the NULL stream) into a
 * graph owned by the context; sgi_views_replay launches it.  The graph keeps
 * the batch's shape: after sgi_views_set with other sizes, capture again
 * (sgi_views_replay fails until then); new buffer addresses or frame data of
 * the same shape are picked up from the device descriptors. */
int sgi_views_capture(sgi_views* views, void* stream);
int sgi_views_replay(sgi_views* views, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SGI_H */
