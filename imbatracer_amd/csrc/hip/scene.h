// Internal definition of a resident scene.
#pragma once

#include <cstddef>

#include "imba_hip.h"

namespace imba {

struct SceneInfo {
    int max_stack = 0;      // deepest traversal stack any ray can need
    bool all_opaque = true; // every mask texel is opaque: the texel test can be skipped
    int treelet_base = 0;   // 4-wide trees: first node of the range that may be staged in LDS
    int treelet_nodes = 0;  // and its length (top levels breadth first; 0 for the BVH2)
};

int validate_scene(const ImbaSceneDesc& d, SceneInfo& info);

// Derives the 4-wide tree of a validated host scene and uploads its nodes and instance records:
// *dev4 = dev with nodes, node_count, instances and root replaced; info4 = info with the
// 4-wide tree's stack bound.  *bytes receives the device bytes allocated.
int upload_wide(const ImbaSceneDesc& host, const ImbaSceneDesc& dev, const SceneInfo& info, ImbaSceneDesc* dev4,
                ImbaSceneDesc* dev4q, SceneInfo* info4, size_t* bytes);
void free_wide(ImbaSceneDesc& dev4, ImbaSceneDesc& dev4q);

}  // namespace imba

struct ImbaScene {
    int device = 0;
    imba::SceneInfo info;
    ImbaSceneDesc counts;  // host view as given at upload (pointers not owned)
    ImbaSceneDesc dev;     // device arrays (owned)
    // the 4-wide tree derived at upload: nodes (ImbaNode4 records) and instance records of its
    // own; the other arrays are those of `dev`
    ImbaSceneDesc dev4;
    imba::SceneInfo info4;
    ImbaSceneDesc dev4q;   // the same tree with quantized nodes (nodes NULL if unavailable)
    int* d_error = nullptr;
    size_t footprint = 0;
};
