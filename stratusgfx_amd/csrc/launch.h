// Host launchers of the GI kernels (each defined next to its kernel).
#pragma once

#include <hip/hip_runtime.h>

#include "gi_device.h"

namespace sgi_launch {

// flags: device scratch of n bytes
hipError_t cullGlobal(const sgi_dev::CsmArgs& csm, const sgi_dev::FrameArgs& fr, const sgi_vpl* in, int n,
                      sgi_vpl* out, int32_t* visible, uint8_t* flags, hipStream_t s);

hipError_t gatherVisibleEntries(const sgi_atlas_entry* all, const int32_t* visible, int maxCount,
                                sgi_atlas_entry* out, hipStream_t s);

hipError_t cullTiles(const sgi_dev::GbufArgs& g, const sgi_dev::FrameArgs& fr, int tileSize, int tilesX, int tilesY,
                     int K, const sgi_vpl* vpls, const int32_t* visible, int remap, sgi_tile_stage1* stage1,
                     int32_t* tiles, hipStream_t s);

hipError_t gather(const sgi_dev::GbufArgs& g, const sgi_dev::FrameArgs& fr, const sgi_dev::CubeArgs& cubes,
                  int format, int tileSize, int tilesX, int tilesY, int K, const sgi_vpl* vpls,
                  const int32_t* visible, const sgi_atlas_entry* entries, const int32_t* tiles, int precision,
                  float* color, float* reservoir, hipStream_t s);

// modes / samples / fp64 as sgi_gather_opts
hipError_t gatherEx(const sgi_dev::GbufArgs& g, const sgi_dev::FrameArgs& fr, const sgi_dev::CubeArgs& cubes,
                    int format, int mode, int samples, int accumulate64, const sgi_vpl* vpls, const int32_t* visible,
                    const sgi_atlas_entry* entries, float* color, float* reservoir, hipStream_t s);

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

}  // namespace sgi_launch
