// Host geometry types used by the BVH builders and layout writers.
//
// The arithmetic follows the reference's core types operation by operation, because the
// builders must reproduce the reference's BVHs bit for bit:
//   float3 ops, cross/dot, comparison-based min/max          core/float3.h:46-128
//   float4 dot (x, y, z, w summed left to right)              core/float4.h:91-93
//   BBox extend/overlap/half_area/empty                       core/bbox.h:18-73
//   Tri::compute_bbox / compute_split / clip_edge             core/tri.h:21-64
//   float3x4 * float4, abs                                    core/float3x4.h:35-50
//   float4x4 invert (cofactor expansion)                      core/float4x4.h:144-191
// This file is compiled with -ffp-contract=off so that no multiply-add is fused.
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace imba {

struct float3 {
    float x, y, z;
    float3() = default;
    explicit float3(float s) : x(s), y(s), z(s) {}
    float3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int a) const { return (&x)[a]; }
    float& operator[](int a) { return (&x)[a]; }
};

inline float3 operator+(const float3& a, const float3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline float3 operator-(const float3& a, const float3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float3 operator*(float s, const float3& b) { return {s * b.x, s * b.y, s * b.z}; }
inline float3 operator*(const float3& a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float3 vabs(const float3& a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
inline float3 cross(const float3& a, const float3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float dot(const float3& a, const float3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length(const float3& a) { return std::sqrt(dot(a, a)); }
inline float3 normalize(const float3& a) { return a * (1.0f / length(a)); }
// Component-wise min/max that return the second operand unless the first is strictly
// smaller/larger (the reference's float3 min/max; NaN and signed zeros behave as there).
inline float3 vmin(const float3& a, const float3& b) {
    return {a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z};
}
inline float3 vmax(const float3& a, const float3& b) {
    return {a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z};
}
// std::min / std::max semantics on floats (second operand only if strictly smaller/larger).
inline float smin(float a, float b) { return (b < a) ? b : a; }
inline float smax(float a, float b) { return (a < b) ? b : a; }

struct float4 {
    float x, y, z, w;
    float4() = default;
    float4(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
    float4(const float3& v, float d) : x(v.x), y(v.y), z(v.z), w(d) {}
    float operator[](int a) const { return (&x)[a]; }
    float& operator[](int a) { return (&x)[a]; }
};
inline float dot4(const float4& a, const float4& b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }

// float -> int conversion with the x86 cvttss2si result for values that do not fit
// (NaN, +-inf, |x| >= 2^31): INT_MIN.  The reference's bin computations cast NaN and
// infinities to int (fast_bvh_builder.h:215, sbvh_builder.h:311-312) and rely on this.
inline int cvt_trunc(float f) {
    if (!(f > -2147483904.0f && f < 2147483648.0f)) return INT32_MIN;
    return (int)f;
}

struct BBox {
    float3 min, max;
    BBox() = default;
    BBox(const float3& lo, const float3& hi) : min(lo), max(hi) {}
    static BBox empty() { return BBox(float3(FLT_MAX), float3(-FLT_MAX)); }
    BBox& extend(const float3& p) { min = vmin(min, p); max = vmax(max, p); return *this; }
    BBox& extend(const BBox& b) { min = vmin(min, b.min); max = vmax(max, b.max); return *this; }
    BBox& overlap(const BBox& b) { min = vmax(min, b.min); max = vmin(max, b.max); return *this; }
    float half_area() const {
        const float3 len = max - min;
        const float kx = smax(len.x, 0.0f);
        const float ky = smax(len.y, 0.0f);
        const float kz = smax(len.z, 0.0f);
        return kx * (ky + kz) + ky * kz;
    }
    bool is_empty() const { return min.x > max.x || min.y > max.y || min.z > max.z; }
};

struct Tri {
    float3 v0, v1, v2;
    Tri() = default;
    Tri(const float3& a, const float3& b, const float3& c) : v0(a), v1(b), v2(c) {}

    void compute_bbox(BBox& bb) const {
        bb.min = vmin(v0, vmin(v1, v2));
        bb.max = vmax(v0, vmax(v1, v2));
    }

    // Bounds of the parts of the triangle on either side of the plane `axis = split`
    // (vertices on the plane belong to the left part).
    void compute_split(BBox& left, BBox& right, int axis, float split) const {
        left = BBox::empty();
        right = BBox::empty();
        const float3 verts[3] = {v0, v1, v2};
        const float3 edges[3] = {v1 - v0, v2 - v1, v0 - v2};
        bool on_left[3];
        for (int i = 0; i < 3; i++) on_left[i] = verts[i][axis] <= split;
        for (int i = 0; i < 3; i++) (on_left[i] ? left : right).extend(verts[i]);
        for (int i = 0; i < 3; i++) {
            const int j = (i + 1) % 3;
            if (on_left[i] != on_left[j]) {
                const float t = (split - verts[i][axis]) / edges[i][axis];
                const float3 p = verts[i] + t * edges[i];
                left.extend(p);
                right.extend(p);
            }
        }
    }
};

struct Mat4 {
    float m[4][4];
};

// Inverse by cofactor expansion in the reference's term order; a singular matrix gives
// the zero matrix.
Mat4 invert(const Mat4& a);

}  // namespace imba
