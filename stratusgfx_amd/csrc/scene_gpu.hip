// Device ray casts of the seeded scene generator (same per-texel code as the
// host variants in scene.cpp).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <string>

#include "scene_internal.h"

namespace {

struct DeviceScene {
    sgi_rt::SceneView view;
};

void freeDeviceMirrors(sgi_scene* s) {
    if (s->dBoxes) (void)hipFree(s->dBoxes);
    if (s->dCellStart) (void)hipFree(s->dCellStart);
    if (s->dCellItems) (void)hipFree(s->dCellItems);
    s->dBoxes = s->dCellStart = s->dCellItems = nullptr;
}

bool hipOk(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    sgi_set_error(std::string(what) + ": " + hipGetErrorString(e));
    return false;
}

// Uploads boxes and the grid once; returns the device-side view.
bool deviceView(sgi_scene* s, sgi_rt::SceneView& v) {
    if (!s->dBoxes) {
        const size_t nb = s->boxes.size() * sizeof(sgi_rt::Box);
        if (!hipOk(hipMalloc(&s->dBoxes, nb), "hipMalloc(boxes)")) return false;
        s->freeDevice = freeDeviceMirrors;
        if (!hipOk(hipMemcpy(s->dBoxes, s->boxes.data(), nb, hipMemcpyHostToDevice), "hipMemcpy(boxes)")) return false;
        if (!s->cellStart.empty()) {
            const size_t ns = s->cellStart.size() * sizeof(int), ni = std::max<size_t>(1, s->cellItems.size()) * sizeof(int);
            if (!hipOk(hipMalloc(&s->dCellStart, ns), "hipMalloc(grid)")) return false;
            if (!hipOk(hipMalloc(&s->dCellItems, ni), "hipMalloc(grid)")) return false;
            if (!hipOk(hipMemcpy(s->dCellStart, s->cellStart.data(), ns, hipMemcpyHostToDevice), "hipMemcpy(grid)")) return false;
            if (!s->cellItems.empty() &&
                !hipOk(hipMemcpy(s->dCellItems, s->cellItems.data(), s->cellItems.size() * sizeof(int),
                                 hipMemcpyHostToDevice), "hipMemcpy(grid)"))
                return false;
        }
    }
    v.boxes = (const sgi_rt::Box*)s->dBoxes;
    v.numAlways = s->numAlways;
    v.numBoxes = (int)s->boxes.size();
    v.grid = s->grid;
    v.grid.cellStart = (const int*)s->dCellStart;
    v.grid.cellItems = (const int*)s->dCellItems;
    if (s->cellStart.empty()) v.grid.nx = 0;
    return true;
}

__global__ void gbufferKernel(sgi_rt::SceneView sv, sgi_rt::CameraRays cam, float* depth, float* normal,
                              float* albedo, float* refl, float* rma, float* velocity, float* id) {
    const long n = (long)cam.width * cam.height;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    sgi_rt::GbufTexel g;
    sgi_rt::shadeGbufferTexel(sv, cam, (int)(p % cam.width), (int)(p / cam.width), g);
    if (depth) depth[p] = g.depth;
    if (normal) for (int k = 0; k < 3; ++k) normal[p * 3 + k] = g.normal[k];
    if (albedo) for (int k = 0; k < 4; ++k) albedo[p * 4 + k] = g.albedo[k];
    if (refl) for (int k = 0; k < 2; ++k) refl[p * 2 + k] = g.refl[k];
    if (rma) for (int k = 0; k < 3; ++k) rma[p * 3 + k] = g.rma[k];
    if (velocity) for (int k = 0; k < 2; ++k) velocity[p * 2 + k] = g.velocity[k];
    if (id) id[p] = g.id;
}

struct CubeJob {
    const sgi_vpl* vpls;    // device, all scene VPLs
    int first, count, res, format, layersPerAtlas;
    void* atlas[SGI_MAX_TOTAL_SHADOW_ATLASES];
};

__global__ void cubeKernel(sgi_rt::SceneView sv, CubeJob job) {
    const long perCube = 6L * job.res * job.res;
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long)job.count * perCube) return;
    const int j = job.first + (int)(q / perCube);
    const long r = q % perCube;
    const int face = (int)(r / ((long)job.res * job.res));
    const int t = (int)((r / job.res) % job.res), i = (int)(r % job.res);
    const sgi_vpl v = job.vpls[j];
    const int atlas = j / job.layersPerAtlas, layer = j % job.layersPerAtlas;
    const size_t off = (((size_t)layer * 6 + face) * job.res + t) * job.res + i;
    if (job.format == SGI_TEXEL_RGBA8) {
        ((uint32_t*)job.atlas[atlas])[off] = sgi_rt::cubeTexelColor(sv, v.position, v.farPlane, face, i, t, job.res);
        return;
    }
    const float val = sgi_rt::cubeTexelValue(sv, v.position, v.farPlane, face, i, t, job.res);
    if (job.format == SGI_TEXEL_F32) ((float*)job.atlas[atlas])[off] = val;
    else if (job.format == SGI_TEXEL_UNORM16) ((uint16_t*)job.atlas[atlas])[off] = sgi_rt::encodeUnorm16(val);
    else ((__half*)job.atlas[atlas])[off] = __float2half_rn(val);
}

__global__ void csmKernel(sgi_rt::SceneView sv, sgi_rt::CsmRays rays, int res, float* depth) {
    const long per = (long)res * res;
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 4 * per) return;
    const int c = (int)(q / per);
    const int j = (int)((q % per) / res), i = (int)(q % res);
    depth[q] = sgi_rt::csmTexelValue(sv, rays, c, i, j, res);
}

}  // namespace

extern "C" {

int sgi_scene_gbuffer_device(sgi_scene* s, float* depth, float* normal, float* albedo, float* refl,
                             float* rma, float* velocity, float* id, void* stream) {
    if (!s || !s->hasView) { sgi_set_error("sgi_scene_gbuffer_device: no view set"); return SGI_E_STATE; }
    sgi_rt::SceneView v;
    if (!deviceView(s, v)) return SGI_E_HIP;
    const long n = (long)s->cam.width * s->cam.height;
    hipLaunchKernelGGL(gbufferKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, v,
                       s->cam, depth, normal, albedo, refl, rma, velocity, id);
    return hipOk(hipGetLastError(), "gbufferKernel") ? SGI_OK : SGI_E_HIP;
}

int sgi_scene_cubes_device(sgi_scene* s, const sgi_cube_atlases* c, int32_t first, int32_t count, void* stream) {
    if (!s || !c) { sgi_set_error("sgi_scene_cubes_device: NULL argument"); return SGI_E_INVALID; }
    if (c->resolution <= 0 ||
        (c->format != SGI_TEXEL_F32 && c->format != SGI_TEXEL_F16 && c->format != SGI_TEXEL_RGBA8 &&
         c->format != SGI_TEXEL_UNORM16) ||
        c->layers_per_atlas <= 0 || c->num_atlases <= 0 || c->num_atlases > SGI_MAX_TOTAL_SHADOW_ATLASES ||
        first < 0 || count < 0 || first + count > (int)s->vpls.size() ||
        (long)(first + count) > (long)c->layers_per_atlas * c->num_atlases) {
        sgi_set_error("sgi_scene_cubes_device: bad atlas description or VPL range");
        return SGI_E_INVALID;
    }
    for (int a = 0; a < c->num_atlases; ++a)
        if (!c->atlas[a]) { sgi_set_error("sgi_scene_cubes_device: NULL atlas"); return SGI_E_INVALID; }
    if (count == 0) return SGI_OK;
    sgi_rt::SceneView v;
    if (!deviceView(s, v)) return SGI_E_HIP;
    sgi_vpl* dv = nullptr;
    const size_t nb = s->vpls.size() * sizeof(sgi_vpl);
    if (!hipOk(hipMalloc(&dv, nb), "hipMalloc(vpls)")) return SGI_E_HIP;
    if (!hipOk(hipMemcpy(dv, s->vpls.data(), nb, hipMemcpyHostToDevice), "hipMemcpy(vpls)")) {
        (void)hipFree(dv);
        return SGI_E_HIP;
    }
    CubeJob job{};
    job.vpls = dv;
    job.first = first;
    job.count = count;
    job.res = c->resolution;
    job.format = c->format;
    job.layersPerAtlas = c->layers_per_atlas;
    for (int a = 0; a < c->num_atlases; ++a) job.atlas[a] = const_cast<void*>(c->atlas[a]);
    const long total = (long)count * 6L * c->resolution * c->resolution;
    hipLaunchKernelGGL(cubeKernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream, v, job);
    const bool ok = hipOk(hipGetLastError(), "cubeKernel") && hipOk(hipStreamSynchronize((hipStream_t)stream), "cubeKernel");
    (void)hipFree(dv);
    return ok ? SGI_OK : SGI_E_HIP;
}

int sgi_scene_csm_device(sgi_scene* s, float* depth, void* stream) {
    if (!s || !s->hasView || !depth) { sgi_set_error("sgi_scene_csm_device: no view set or NULL"); return SGI_E_STATE; }
    sgi_rt::SceneView v;
    if (!deviceView(s, v)) return SGI_E_HIP;
    const long n = 4L * s->csmResolution * s->csmResolution;
    hipLaunchKernelGGL(csmKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, v, s->csm,
                       s->csmResolution, depth);
    return hipOk(hipGetLastError(), "csmKernel") ? SGI_OK : SGI_E_HIP;
}

}  // extern "C"
