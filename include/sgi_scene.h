/*
 * sgi_scene.h — C ABI of the seeded synthetic input generator.
 *
 * No counterpart in the reference's GI pass: the reference rasterises these
 * inputs with its geometry, VPL-shadow and cascade passes
 * (pbr_geometry_pass.fs, shadowVpl.fs, csm.fs), which sit outside the pass.
 * This generator ray-casts an analytic scene instead: a Cornell box (5 walls,
 * red / green sides, 2 boxes rotated about Y) or a procedural city block (a
 * ground slab plus N random axis-aligned boxes), lit by one directional sun.
 * It writes the pass inputs in the layouts of sgi.h.  Host (_host) and device
 * (_device) variants produce the same data from the same code.
 */
#ifndef SGI_SCENE_H
#define SGI_SCENE_H

#include "sgi.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { SGI_SCENE_CORNELL = 0, SGI_SCENE_CITY = 1 };

typedef struct sgi_scene_desc {
    int32_t kind;                /* SGI_SCENE_CORNELL or SGI_SCENE_CITY */
    uint64_t seed;
    int32_t num_vpls;            /* VPLs sampled on sun-lit surfaces, <= 10000 */
    int32_t num_city_boxes;      /* city only (2000) */
    float city_half_extent;      /* city only: boxes inside [-e, e]^2 in XZ (100 m) */
    float vpl_color_multiplier;  /* VPL colour = albedo * irradiance * this */
} sgi_scene_desc;

typedef struct sgi_view {
    int32_t width;
    int32_t height;
    float eye[3];
    float yaw_deg;               /* 0 looks down -Z, positive turns toward +X */
    float pitch_deg;             /* positive looks up */
    float fovy_deg;              /* 75 in the reference (StratusEngine.cpp:213) */
    float znear;                 /* 1 */
    float zfar;                  /* 1000 */
    int32_t csm_resolution;      /* cascade edge in texels */
} sgi_view;

typedef struct sgi_scene sgi_scene;

int sgi_scene_create(const sgi_scene_desc* desc, sgi_scene** out);
void sgi_scene_destroy(sgi_scene* scene);
int sgi_scene_num_boxes(const sgi_scene* scene);

/* The scene's boxes (host memory, 16 floats each): centre xyz, cos(yaw),
 * half extents xyz, sin(yaw), albedo rgb, roughness, metallic, mesh id, 0, 0. */
int sgi_scene_boxes(const sgi_scene* scene, float* out);

/* A view of the scene that frames it (camera at the Cornell box mouth, or over
 * the city block). */
int sgi_scene_default_view(const sgi_scene* scene, int32_t width, int32_t height, sgi_view* out);

/* Makes `view` the scene's current view (prev = previous frame's view for the
 * G-buffer velocity; NULL = same view) and fills the frame uniforms and the
 * cascade parameters (csm->depth is left NULL). */
int sgi_scene_set_view(sgi_scene* scene, const sgi_view* view, const sgi_view* prev,
                       sgi_frame* frame, sgi_csm* csm);

/* The sampled VPLs (host memory): records with position, colour, specular
 * position, radius = farPlane = 500, intensity 1, and the cube atlas entry of
 * VPL j = { j / layers_per_atlas, j % layers_per_atlas }. */
int sgi_scene_vpls(const sgi_scene* scene, int32_t layers_per_atlas,
                   sgi_vpl* vpls, sgi_atlas_entry* entries);

/* G-buffer of the current view; any output may be NULL. */
int sgi_scene_gbuffer_host(const sgi_scene* scene, float* depth, float* normal, float* albedo,
                           float* baseReflectivity, float* roughnessMetallicAmbient,
                           float* velocity, float* id);
int sgi_scene_gbuffer_device(sgi_scene* scene, float* depth, float* normal, float* albedo,
                             float* baseReflectivity, float* roughnessMetallicAmbient,
                             float* velocity, float* id, void* stream);

/* Cube maps of VPLs [first, first + count) into the atlases of `cubes`
 * (resolution, format, layers_per_atlas, num_atlases and atlas[] must be set;
 * for _host the atlas pointers are host memory): distance / farPlane for the
 * F32 and F16 formats (the depth attachment shadowVpl.fs:29-32 writes), the
 * nearest surface's albedo as UNORM8 RGB for SGI_TEXEL_RGBA8 (its colour
 * attachment, shadowVpl.fs:37; 0 where nothing lies within farPlane). */
int sgi_scene_cubes_host(const sgi_scene* scene, const sgi_cube_atlases* cubes,
                         int32_t first, int32_t count);
int sgi_scene_cubes_device(sgi_scene* scene, const sgi_cube_atlases* cubes,
                           int32_t first, int32_t count, void* stream);

/* Sun cascades of the current view: depth [4][res][res]. */
int sgi_scene_csm_host(const sgi_scene* scene, float* depth);
int sgi_scene_csm_device(sgi_scene* scene, float* depth, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SGI_SCENE_H */
