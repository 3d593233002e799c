// C ABI of the batched views (include/sgi.h, "batched views"): per-view
// descriptors on the device, one launch per stage over every view, and the
// optional captured graph of one frame.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/sgi.h"
#include "../../include/sgi_kpcn.h"
#include "launch.h"

void sgi_set_error(const std::string& msg);

struct sgi_views {
    sgi_views_config cfg{};
    int device = 0;
    int numViews = 0, totalTiles = 0, totalTiles16 = 0, totalBlocks = 0, numVpls = 0, maxEntries = 0;
    const sgi_vpl* vpls = nullptr;
    const sgi_atlas_entry* shadowIndices = nullptr;
    sgi_dev::CubeArgs cubes{};
    int cubeFormat = SGI_TEXEL_F32;
    bool hasBatch = false;
    // device scratch: per view the culling outputs, per pixel the denoiser planes
    sgi_batch::ViewDesc* descs = nullptr;      // [max_views]
    sgi_batch::ViewDesc* hostDescs = nullptr;  // pinned staging of the descriptors
    sgi_vpl* vplUpdated = nullptr;             // [max_views][4480]
    int32_t* visible = nullptr;                // [max_views][1 + 4480]
    sgi_atlas_entry* entries = nullptr;        // [max_views][4480]
    uint8_t* visFlags = nullptr;               // [max_views][10000]
    float4* planes = nullptr;                  // [3][max_pixels]
    // the learned form (samples > 0): per-view sampled gather + learned denoise
    int samples = 0, rejectionLimit = -1;
    sgi_kpcn_net net{};
    float threshold = 0.0f;
    void* prep = nullptr;        // the sampled gather's packed records (views run one after another)
    void* kws = nullptr;         // the network's workspace, for the largest view
    size_t kwsBytes = 0;
    std::vector<sgi_batch::ViewDesc> host;  // the descriptors on the host (per-view launches)
    // the captured frame
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // what the captured launches baked in (grid sizes and by-value arguments)
    struct Shape {
        int numViews, totalTiles, totalTiles16, totalBlocks, numVpls, cubeFormat;
        const sgi_vpl* vpls;
        const sgi_atlas_entry* shadowIndices;
        sgi_dev::CubeArgs cubes;
        int samples, rejectionLimit;
        sgi_kpcn_net net;
        float threshold;
        const void* kws;
    } captured{};
    bool hasGraph = false;
};

namespace {

int fail(int code, const std::string& msg) {
    sgi_set_error(msg);
    return code;
}

int hipFail(hipError_t e, const char* what) { return fail(SGI_E_HIP, std::string(what) + ": " + hipGetErrorString(e)); }

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

sgi_dev::CsmArgs csmArgs(const sgi_csm& c) {
    sgi_dev::CsmArgs a;
    a.depth = c.depth;
    a.res = c.resolution;
    for (int i = 0; i < SGI_NUM_CASCADES; ++i)
        for (int k = 0; k < 16; ++k) a.projViews[i][k] = c.cascadeProjViews[i][k];
    for (int i = 0; i < SGI_NUM_CASCADES - 1; ++i)
        for (int k = 0; k < 4; ++k) a.planes[i][k] = c.cascadePlanes[i][k];
    a.bias = c.depthBias / (c.zfar - c.znear);  // pbr.glsl:116
    return a;
}

void releaseGraph(sgi_views* v) {
    if (v->exec) (void)hipGraphExecDestroy(v->exec);
    if (v->graph) (void)hipGraphDestroy(v->graph);
    v->exec = nullptr;
    v->graph = nullptr;
    v->hasGraph = false;
}

sgi_views::Shape shapeOf(const sgi_views* v) {
    sgi_views::Shape s{};
    s.numViews = v->numViews;
    s.totalTiles = v->totalTiles;
    s.totalTiles16 = v->totalTiles16;
    s.totalBlocks = v->totalBlocks;
    s.numVpls = v->numVpls;
    s.cubeFormat = v->cubeFormat;
    s.vpls = v->vpls;
    s.shadowIndices = v->shadowIndices;
    s.cubes = v->cubes;
    s.samples = v->samples;
    s.rejectionLimit = v->rejectionLimit;
    s.net = v->net;
    s.threshold = v->threshold;
    s.kws = v->kws;
    return s;
}

bool sameShape(const sgi_views::Shape& a, const sgi_views::Shape& b) {
    bool same = a.numViews == b.numViews && a.totalTiles == b.totalTiles && a.totalTiles16 == b.totalTiles16 &&
                a.totalBlocks == b.totalBlocks && a.numVpls == b.numVpls && a.cubeFormat == b.cubeFormat &&
                a.vpls == b.vpls && a.shadowIndices == b.shadowIndices && a.cubes.numAtlases == b.cubes.numAtlases &&
                a.cubes.res == b.cubes.res && a.cubes.layers == b.cubes.layers;
    for (int i = 0; same && i < SGI_MAX_TOTAL_SHADOW_ATLASES; ++i) same = a.cubes.atlas[i] == b.cubes.atlas[i];
    same = same && a.samples == b.samples && a.rejectionLimit == b.rejectionLimit && a.kws == b.kws &&
           a.net.weights == b.net.weights && a.net.biases == b.net.biases &&
           a.net.weights_fp8 == b.net.weights_fp8 && a.net.epilogue == b.net.epilogue &&
           std::memcmp(&a.threshold, &b.threshold, sizeof(float)) == 0;
    return same;
}

// The network's workspace for the batch's largest view (outside any capture).
hipError_t sizeLearnedWorkspace(sgi_views* v) {
    size_t need = 0;
    for (const sgi_batch::ViewDesc& d : v->host) need = std::max(need, sgi_kpcn_workspace_bytes(d.g.H, d.g.W));
    if (need <= v->kwsBytes) return hipSuccess;
    if (v->kws) (void)hipFree(v->kws);
    v->kws = nullptr;
    v->kwsBytes = 0;
    const hipError_t e = hipMalloc(&v->kws, need);
    if (e == hipSuccess) v->kwsBytes = need;
    return e;
}

hipError_t runStages(sgi_views* v, hipStream_t s) {
    const sgi_views_config& c = v->cfg;
    hipError_t e = sgi_launch::cullGlobalViews(v->descs, v->numViews, v->vpls, v->numVpls, s);
    if (e == hipSuccess)
        e = sgi_launch::gatherVisibleEntriesViews(v->descs, v->numViews, v->shadowIndices, v->maxEntries, s);
    if (e == hipSuccess) e = sgi_launch::directLightViews(v->descs, v->numViews, v->totalTiles16, s);
    if (v->samples > 0) {
        // the learned form: each view's sampled gather and network in turn (sizes differ per view)
        for (int i = 0; i < v->numViews && e == hipSuccess; ++i) {
            const sgi_batch::ViewDesc& d = v->host[(size_t)i];
            e = sgi_launch::gatherEx(d.g, d.fr, v->cubes, v->cubeFormat, SGI_GATHER_SAMPLES, v->samples,
                                     v->rejectionLimit, 0, 0, d.vplUpdated, d.visible, d.entries, v->prep, d.color,
                                     d.reservoir, s);
            if (e == hipSuccess)
                e = sgi_launch::denoiseLearned(d.g, d.g, d.fr, d.visible, d.direct, d.color, d.reservoir, d.prevGi,
                                               d.prevHistory, v->net, v->threshold, v->kws, d.combined, d.giOut,
                                               d.reservoirOut, d.historyOut, nullptr,
                                               sgi_launch::DenoiseBuffers{d.resA, d.resB, d.normals}, nullptr, s);
        }
        return e;
    }
    if (e == hipSuccess)
        e = sgi_launch::cullTilesViews(v->descs, v->numViews, v->totalTiles, c.tile_size, c.max_vpls_per_tile, s);
    if (e == hipSuccess)
        e = sgi_launch::gatherViews(v->descs, v->numViews, v->totalTiles, c.tile_size, c.max_vpls_per_tile, v->cubes,
                                    v->cubeFormat, s);
    if (e == hipSuccess) e = sgi_launch::denoiseViews(v->descs, v->numViews, v->totalBlocks, s);
    return e;
}

}  // namespace

extern "C" {

int sgi_views_create(const sgi_views_config* config, sgi_views** out) {
    if (!config || !out) return fail(SGI_E_INVALID, "sgi_views_create: NULL argument");
    *out = nullptr;
    const sgi_views_config& c = *config;
    if (c.max_views < 1 || c.max_views > 64) return fail(SGI_E_INVALID, "sgi_views_create: max_views must be 1..64");
    if (c.max_pixels < 1 || c.max_pixels > ((int64_t)1 << 31))
        return fail(SGI_E_INVALID, "sgi_views_create: max_pixels must be within 1..2^31");
    if (c.tile_size != 8 && c.tile_size != 16 && c.tile_size != 32)
        return fail(SGI_E_INVALID, "sgi_views_create: tile_size must be 8, 16 or 32");
    if (c.max_vpls_per_tile < 1 || c.max_vpls_per_tile > SGI_MAX_VPLS_PER_TILE_LIMIT)
        return fail(SGI_E_INVALID, "sgi_views_create: max_vpls_per_tile must be within 1..64");
    sgi_views* v = new (std::nothrow) sgi_views();
    if (!v) return fail(SGI_E_INVALID, "sgi_views_create: out of host memory");
    v->cfg = c;
    const size_t V = (size_t)c.max_views, cap = SGI_MAX_TOTAL_VPLS_PER_FRAME;
    hipError_t e = hipGetDevice(&v->device);
    if (e == hipSuccess) e = hipMalloc(&v->descs, V * sizeof(sgi_batch::ViewDesc));
    if (e == hipSuccess) e = hipHostMalloc((void**)&v->hostDescs, V * sizeof(sgi_batch::ViewDesc), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&v->vplUpdated, V * cap * sizeof(sgi_vpl));
    if (e == hipSuccess) e = hipMalloc(&v->visible, V * (1 + cap) * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&v->entries, V * cap * sizeof(sgi_atlas_entry));
    if (e == hipSuccess) e = hipMalloc(&v->visFlags, V * SGI_MAX_TOTAL_VPLS_BEFORE_CULLING);
    if (e == hipSuccess) e = hipMalloc(&v->planes, 3 * (size_t)c.max_pixels * sizeof(float4));
    if (e != hipSuccess) {
        sgi_views_destroy(v);
        return hipFail(e, "sgi_views_create: workspace allocation");
    }
    *out = v;
    return SGI_OK;
}

void sgi_views_destroy(sgi_views* v) {
    if (!v) return;
    releaseGraph(v);
    if (v->descs) (void)hipFree(v->descs);
    if (v->hostDescs) (void)hipHostFree(v->hostDescs);
    if (v->vplUpdated) (void)hipFree(v->vplUpdated);
    if (v->visible) (void)hipFree(v->visible);
    if (v->entries) (void)hipFree(v->entries);
    if (v->visFlags) (void)hipFree(v->visFlags);
    if (v->planes) (void)hipFree(v->planes);
    if (v->prep) (void)hipFree(v->prep);
    if (v->kws) (void)hipFree(v->kws);
    delete v;
}

int sgi_views_set(sgi_views* v, int32_t numViews, const sgi_view_io* io, const sgi_vpl* vpls, int32_t numVpls,
                  const sgi_atlas_entry* shadowIndices, const sgi_cube_atlases* cubes, void* stream) {
    if (!v || !io || !cubes) return fail(SGI_E_INVALID, "sgi_views_set: NULL argument");
    if (numViews < 1 || numViews > v->cfg.max_views)
        return fail(SGI_E_INVALID, "sgi_views_set: numViews must be within 1..max_views");
    if (numVpls < 0 || numVpls > SGI_MAX_TOTAL_VPLS_BEFORE_CULLING)
        return fail(SGI_E_INVALID, "sgi_views_set: numVpls must be within 0..10000");
    if (numVpls > 0 && (!vpls || !shadowIndices)) return fail(SGI_E_INVALID, "sgi_views_set: NULL VPL buffer");
    if (numVpls > 0 && ((uintptr_t)vpls & 15u) != 0)
        return fail(SGI_E_INVALID, "sgi_views_set: VPL records must be 16-byte aligned");
    if (cubes->resolution <= 0 || cubes->resolution > 4096 ||
        (cubes->format != SGI_TEXEL_F32 && cubes->format != SGI_TEXEL_F16 && cubes->format != SGI_TEXEL_UNORM16) ||
        cubes->layers_per_atlas <= 0 ||
        cubes->layers_per_atlas > SGI_MAX_TOTAL_SHADOWS_PER_ATLAS || cubes->num_atlases <= 0 ||
        cubes->num_atlases > SGI_MAX_TOTAL_SHADOW_ATLASES)
        return fail(SGI_E_INVALID, "sgi_views_set: bad cube atlas description");
    sgi_dev::CubeArgs ca{};
    for (int a = 0; a < cubes->num_atlases; ++a) {
        if (!cubes->atlas[a]) return fail(SGI_E_INVALID, "sgi_views_set: NULL cube atlas");
        ca.atlas[a] = cubes->atlas[a];
    }
    for (int a = cubes->num_atlases; a < SGI_MAX_TOTAL_SHADOW_ATLASES; ++a) ca.atlas[a] = cubes->atlas[0];
    ca.numAtlases = cubes->num_atlases;
    ca.res = cubes->resolution;
    ca.layers = cubes->layers_per_atlas;
    ca.layerStride = 6u * (uint32_t)cubes->resolution * (uint32_t)cubes->resolution;

    const int ts = v->cfg.tile_size, cap = SGI_MAX_TOTAL_VPLS_PER_FRAME;
    int64_t pix = 0;
    int tiles = 0, tiles16 = 0, blocks = 0;
    for (int i = 0; i < numViews; ++i) {
        const sgi_view_io& x = io[i];
        const sgi_gbuffer& g = x.gbuffer;
        if (g.width <= 0 || g.height <= 0 || g.width > 16384 || g.height > 16384)
            return fail(SGI_E_INVALID, "sgi_views_set: view size must be within 1..16384");
        if (g.packed) {
            const sgi_gbuffer_packed& k = *g.packed;
            if (k.width != g.width || k.height != g.height || !g.depth || !k.normals || !k.albedo ||
                !k.baseReflectivity || !k.roughnessMetallicAmbient)
                return fail(SGI_E_INVALID, "sgi_views_set: a view's packed attachments are NULL or of another size");
        } else if (!g.depth || !g.normal || !g.albedo || !g.baseReflectivity || !g.roughnessMetallicAmbient) {
            return fail(SGI_E_INVALID, "sgi_views_set: a view's G-buffer plane is NULL");
        }
        if (!x.tiles || !x.direct || !x.color || !x.reservoir || !x.combined || !x.giOut || !x.reservoirOut ||
            !x.historyOut)
            return fail(SGI_E_INVALID, "sgi_views_set: a view's buffer is NULL");
        if (!aligned16(x.reservoir) || !aligned16(x.reservoirOut))
            return fail(SGI_E_INVALID, "sgi_views_set: reservoirs must be 16-byte aligned");
        if (!(x.frame.framesPerSecond > 0.0f)) return fail(SGI_E_INVALID, "sgi_views_set: framesPerSecond must be > 0");
        if (x.csm.resolution <= 0 || x.csm.resolution > 16384 || !x.csm.depth)
            return fail(SGI_E_INVALID, "sgi_views_set: a view's cascades are missing");
        sgi_batch::ViewDesc& d = v->hostDescs[i];
        d.g = sgi_dev::gbufArgs(g);
        d.fr = sgi_dev::frameArgs(x.frame);
        d.csm = csmArgs(x.csm);
        d.tilesX = (g.width + ts - 1) / ts;
        d.tilesY = (g.height + ts - 1) / ts;
        d.tileStart = tiles;
        d.dtileStart = tiles16;
        d.dblockStart = blocks;
        tiles += d.tilesX * d.tilesY;
        tiles16 += ((g.width + 15) / 16) * ((g.height + 15) / 16);
        blocks += ((g.width + 63) / 64) * ((g.height + 3) / 4);
        d.vplUpdated = v->vplUpdated + (size_t)i * cap;
        d.visible = v->visible + (size_t)i * (1 + cap);
        d.entries = v->entries + (size_t)i * cap;
        d.visFlags = v->visFlags + (size_t)i * SGI_MAX_TOTAL_VPLS_BEFORE_CULLING;
        d.tiles = x.tiles;
        d.direct = x.direct;
        d.color = x.color;
        d.reservoir = x.reservoir;
        d.prevGi = x.prevGi;
        d.prevHistory = x.prevHistory;
        d.combined = x.combined;
        d.giOut = x.giOut;
        d.reservoirOut = x.reservoirOut;
        d.historyOut = x.historyOut;
        d.resA = v->planes + pix;
        d.resB = v->planes + v->cfg.max_pixels + pix;
        d.normals = v->planes + 2 * v->cfg.max_pixels + pix;
        pix += (int64_t)g.width * g.height;
        if (pix > v->cfg.max_pixels) return fail(SGI_E_INVALID, "sgi_views_set: the views exceed max_pixels");
    }
    hipError_t e = hipMemcpyAsync(v->descs, v->hostDescs, (size_t)numViews * sizeof(sgi_batch::ViewDesc),
                                  hipMemcpyHostToDevice, (hipStream_t)stream);
    // the staging buffer is rewritten by the next call, so the copy must have completed
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return hipFail(e, "sgi_views_set");
    try {
        v->host.assign(v->hostDescs, v->hostDescs + numViews);
    } catch (const std::bad_alloc&) {
        return fail(SGI_E_INVALID, "sgi_views_set: out of host memory");
    }
    if (v->samples > 0 && (e = sizeLearnedWorkspace(v)) != hipSuccess)
        return hipFail(e, "sgi_views_set: learned workspace");
    v->numViews = numViews;
    v->totalTiles = tiles;
    v->totalTiles16 = tiles16;
    v->totalBlocks = blocks;
    v->numVpls = numVpls;
    v->maxEntries = numVpls < cap ? numVpls : cap;
    v->vpls = vpls;
    v->shadowIndices = shadowIndices;
    v->cubes = ca;
    v->cubeFormat = cubes->format;
    v->hasBatch = true;
    return SGI_OK;
}

int sgi_views_run(sgi_views* v, void* stream) {
    if (!v) return fail(SGI_E_INVALID, "sgi_views_run: NULL context");
    if (!v->hasBatch) return fail(SGI_E_STATE, "sgi_views_run: sgi_views_set first");
    const hipError_t e = runStages(v, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_views_run");
}

int sgi_views_capture(sgi_views* v, void* stream) {
    if (!v) return fail(SGI_E_INVALID, "sgi_views_capture: NULL context");
    if (!v->hasBatch) return fail(SGI_E_STATE, "sgi_views_capture: sgi_views_set first");
    if (!stream) return fail(SGI_E_INVALID, "sgi_views_capture: capture needs a stream other than the NULL stream");
    releaseGraph(v);
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hipFail(e, "sgi_views_capture: begin");
    const hipError_t er = runStages(v, s);
    e = hipStreamEndCapture(s, &v->graph);
    if (er != hipSuccess) {
        releaseGraph(v);
        return hipFail(er, "sgi_views_capture: stages");
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&v->exec, v->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        releaseGraph(v);
        return hipFail(e, "sgi_views_capture: end / instantiate");
    }
    v->captured = shapeOf(v);
    v->hasGraph = true;
    return SGI_OK;
}

int sgi_views_replay(sgi_views* v, void* stream) {
    if (!v) return fail(SGI_E_INVALID, "sgi_views_replay: NULL context");
    if (!v->hasGraph) return fail(SGI_E_STATE, "sgi_views_replay: sgi_views_capture first");
    if (!sameShape(v->captured, shapeOf(v)))
        return fail(SGI_E_STATE, "sgi_views_replay: the batch changed shape since the capture");
    const hipError_t e = hipGraphLaunch(v->exec, (hipStream_t)stream);
    return e == hipSuccess ? SGI_OK : hipFail(e, "sgi_views_replay");
}

int sgi_views_set_learned(sgi_views* v, int32_t samples, int32_t rejectionLimit, const sgi_kpcn_net* net,
                          float threshold) {
    if (!v) return fail(SGI_E_INVALID, "sgi_views_set_learned: NULL context");
    if (samples == 0) {
        v->samples = 0;
        return SGI_OK;
    }
    if (!v->hasBatch) return fail(SGI_E_STATE, "sgi_views_set_learned: sgi_views_set first");
    if (samples < 1 || samples > 4096) return fail(SGI_E_INVALID, "sgi_views_set_learned: samples must be within 0..4096");
    if (!net || !net->weights || !net->biases || ((uintptr_t)net->weights & 15u) != 0)
        return fail(SGI_E_INVALID, "sgi_views_set_learned: network weights missing or unaligned");
    if (net->weights_fp8 && (!net->epilogue || ((uintptr_t)net->weights_fp8 & 15u) != 0))
        return fail(SGI_E_INVALID, "sgi_views_set_learned: FP8 weights without their epilogue, or unaligned");
    if (std::isnan(threshold)) return fail(SGI_E_INVALID, "sgi_views_set_learned: threshold is NaN");
    hipError_t e = hipSuccess;
    if (!v->prep) e = hipMalloc(&v->prep, (size_t)SGI_MAX_TOTAL_VPLS_PER_FRAME * sgi_launch::packedVplBytes());
    const int previous = v->samples;
    v->samples = samples;
    if (e == hipSuccess) e = sizeLearnedWorkspace(v);
    if (e != hipSuccess) {
        v->samples = previous;
        return hipFail(e, "sgi_views_set_learned: workspace allocation");
    }
    v->rejectionLimit = rejectionLimit;
    v->net = *net;
    v->threshold = threshold;
    return SGI_OK;
}

}  // extern "C"
