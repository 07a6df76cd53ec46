// Four-wide traversal structure derived from a two-level BVH2 scene (host side).
#pragma once

#include <vector>

#include "imba_hip.h"

namespace imba {

struct WideScene {
    std::vector<ImbaNode4> nodes;               // mesh trees, then the top level
    std::vector<ImbaInstanceNode> instances;    // `next` = 4-wide root of the mesh
    int root = 0;
    int max_stack = 0;                          // stack entries a ray can need
};

// `d` must have passed validate_scene (links in range, children after parents).
int collapse_scene4(const ImbaSceneDesc& d, WideScene& out);

}  // namespace imba
