// Batched views: one launch per stage over the work items of every view.
//
// Each view's arguments live in a device array of ViewDesc; a launch's blocks
// are numbered over the concatenation of all views' work items (tiles or
// pixel blocks), so a block finds its view with a short scan of the
// per-view first-item offsets (uniform per block, scalar loads) and then runs
// the same per-tile / per-block body as the single-view kernel.
#pragma once

#include "gi_device.h"

namespace sgi_batch {

struct ViewDesc {
    sgi_dev::GbufArgs g;
    sgi_dev::FrameArgs fr;
    sgi_dev::CsmArgs csm;
    int tilesX, tilesY;  // culling / gather tiles at the batch's tile size
    int tileStart;       // first batch tile of this view
    int dtileStart;      // first batch 16x16 tile (direct light)
    int dblockStart;     // first batch 64x4 block (denoiser)
    sgi_vpl* vplUpdated;
    int32_t* visible;
    sgi_atlas_entry* entries;
    uint8_t* visFlags;
    int32_t* tiles;
    float* direct;
    float* color;
    float* reservoir;
    const float* prevGi;
    const float* prevHistory;
    float* combined;
    float* giOut;
    float* reservoirOut;
    float* historyOut;
    float4* resA;
    float4* resB;
    float4* normals;
};

enum Offset { kTiles = 0, kDirectTiles = 1, kDenoiseBlocks = 2 };

// The view whose work items contain batch item `item` (views are in item order).
template <int OFF>
__device__ __forceinline__ int findView(const ViewDesc* __restrict__ v, int n, int item) {
    int i = 0;
    for (int k = 1; k < n; ++k) {
        const int start = OFF == kTiles ? v[k].tileStart : (OFF == kDirectTiles ? v[k].dtileStart : v[k].dblockStart);
        if (start <= item) i = k;
    }
    return i;
}

}  // namespace sgi_batch
