// Host-side state of a generated scene, shared by scene.cpp (host build and
// host ray casts) and scene_gpu.hip (device ray casts).
#pragma once

#include <string>
#include <vector>

#include "../../include/sgi_scene.h"
#include "scene_rt.h"

struct sgi_scene {
    sgi_scene_desc desc;
    std::vector<sgi_rt::Box> boxes;   // [0, numAlways) always tested, rest gridded
    int numAlways = 0;
    std::vector<int> cellStart, cellItems;
    sgi_rt::Grid grid{};              // pointers patched per use (host / device)
    float sunDir[3];                  // toward the sun
    float sunColor[3];                // luminance
    float boundsMin[3], boundsMax[3];
    std::vector<sgi_vpl> vpls;

    // current view
    bool hasView = false;
    sgi_view view{};
    sgi_rt::CameraRays cam{};
    sgi_rt::CsmRays csm{};
    int csmResolution = 0;

    // device mirrors (allocated on first device call, freed by destroy)
    void* dBoxes = nullptr;
    void* dCellStart = nullptr;
    void* dCellItems = nullptr;
    void (*freeDevice)(sgi_scene*) = nullptr;

    sgi_rt::SceneView hostView() const {
        sgi_rt::SceneView v;
        v.boxes = boxes.data();
        v.numAlways = numAlways;
        v.numBoxes = (int)boxes.size();
        v.grid = grid;
        v.grid.cellStart = cellStart.empty() ? nullptr : cellStart.data();
        v.grid.cellItems = cellItems.empty() ? nullptr : cellItems.data();
        return v;
    }
};

// Thread-local error message shared with the pass library.
void sgi_set_error(const std::string& msg);
