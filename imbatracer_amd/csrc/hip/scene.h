// Internal definition of a resident scene.
#pragma once

#include <cstddef>

#include "imba_hip.h"

namespace imba {

struct SceneInfo {
    int max_stack = 0;      // deepest traversal stack any ray can need
    bool all_opaque = true; // every mask texel is opaque: the texel test can be skipped
};

int validate_scene(const ImbaSceneDesc& d, SceneInfo& info);

}  // namespace imba

struct ImbaScene {
    int device = 0;
    imba::SceneInfo info;
    ImbaSceneDesc counts;  // host view as given at upload (pointers not owned)
    ImbaSceneDesc dev;     // device arrays (owned)
    int* d_error = nullptr;
    size_t footprint = 0;
};
