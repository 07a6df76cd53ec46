// Compile-time check that the reference-named types of traversal_gpu.h and the Imba types of
// imba_hip.h describe the same bytes (the FFI functions are defined on the Imba types).
#include <cstddef>

#include "imba_hip.h"

namespace ref_abi {
#include "traversal_gpu.h"
}

#define SAME_FIELD(A, B, f) static_assert(offsetof(A, f) == offsetof(B, f), #A "." #f)

static_assert(sizeof(ref_abi::Vec4) == sizeof(ImbaVec4) && sizeof(ImbaVec4) == 16, "Vec4");
static_assert(sizeof(ref_abi::Vec2) == sizeof(ImbaVec2) && sizeof(ImbaVec2) == 8, "Vec2");
static_assert(sizeof(ref_abi::Ray) == sizeof(ImbaRay) && sizeof(ImbaRay) == 32, "Ray");
static_assert(sizeof(ref_abi::Hit) == sizeof(ImbaHit) && sizeof(ImbaHit) == 16, "Hit");
static_assert(sizeof(ref_abi::Node) == sizeof(ImbaNode) && sizeof(ImbaNode) == 64, "Node");
static_assert(sizeof(ref_abi::InstanceNode) == sizeof(ImbaInstanceNode) && sizeof(ImbaInstanceNode) == 64,
              "InstanceNode");
static_assert(sizeof(ref_abi::TransparencyMask) == sizeof(ImbaTransparencyMask) && sizeof(ImbaTransparencyMask) == 12,
              "TransparencyMask");
SAME_FIELD(ref_abi::Node, ImbaNode, left_bb);
SAME_FIELD(ref_abi::Node, ImbaNode, right_bb);
SAME_FIELD(ref_abi::Node, ImbaNode, left);
SAME_FIELD(ref_abi::Node, ImbaNode, right);
SAME_FIELD(ref_abi::BBox, ImbaBBox, hi_x);
SAME_FIELD(ref_abi::BBox, ImbaBBox, lo_z);
SAME_FIELD(ref_abi::InstanceNode, ImbaInstanceNode, id);
SAME_FIELD(ref_abi::InstanceNode, ImbaInstanceNode, next);
SAME_FIELD(ref_abi::InstanceNode, ImbaInstanceNode, pad);
SAME_FIELD(ref_abi::Ray, ImbaRay, dir);
SAME_FIELD(ref_abi::Hit, ImbaHit, tmax);
SAME_FIELD(ref_abi::Hit, ImbaHit, u);
SAME_FIELD(ref_abi::TransparencyMask, ImbaTransparencyMask, offset);
