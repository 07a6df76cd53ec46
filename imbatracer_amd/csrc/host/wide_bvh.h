// Four-wide traversal structure derived from a two-level BVH2 scene (host side).
#pragma once

#include <cstdint>
#include <vector>

#include "imba_hip.h"

namespace imba {

// A 4-wide node with child boxes quantized to 8 bits relative to the node's own box (64 B):
// origin (3 floats), per-axis power-of-two scale exponents and the child count, the four
// child links, then per axis the children's low and high bytes (q_lo[4], q_hi[4]).  A child
// box is [fma(q_lo, 2^e, origin), fma(q_hi, 2^e, origin)] per axis, rounded outward, so it
// always contains the exact box.
struct QNode4 {
    float origin[3];
    int8_t exp[3];
    uint8_t count;
    int children[4];
    uint8_t lo_x[4], hi_x[4], lo_y[4], hi_y[4];
    uint8_t lo_z[4], hi_z[4];
    uint32_t pad[2];
};
static_assert(sizeof(QNode4) == 64, "QNode4 layout");

struct WideScene {
    std::vector<ImbaNode4> nodes;               // mesh trees, then the top level
    std::vector<QNode4> qnodes;                 // the same tree, quantized (empty if a box is not finite)
    std::vector<ImbaInstanceNode> instances;    // `next` = 4-wide root of the mesh
    int root = 0;
    int max_stack = 0;                          // stack entries a ray can need
    // Nodes [treelet_base, treelet_base + treelet_nodes) are the first levels, breadth first,
    // of the tree every ray enters first: the top level, or the only mesh when the top level
    // reaches a single instance.  Kernels may stage them in LDS.
    int treelet_base = 0;
    int treelet_nodes = 0;
};

// Nodes of each tree laid out breadth first before the rest of the tree goes depth first.
constexpr int kTreeletMaxNodes = 256;

// `d` must have passed validate_scene (links in range, children after parents).
int collapse_scene4(const ImbaSceneDesc& d, WideScene& out);

// Quantizes every node of `nodes` (filled into w.qnodes); false if a box is not finite.
bool quantize_nodes(WideScene& w);

}  // namespace imba
