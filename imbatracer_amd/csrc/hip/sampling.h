// Random numbers, per-sample seeding and direction sampling shared by the ray producers.
//
// Rng is the MWC64X generator of render/random.h:15-47 (next() returns x ^ c, random_float()
// divides by 2^32).  seeded() is the per-sample seeding of PixelRayGen / LightRayGen
// (render/ray_gen/ray_gen.h:63-70 and :138-145): four rounds of seed = 33 * seed ^ index, then
// (seed % 5) + 16 + (pixel or index) % 5 discarded draws.  frame() is local_coordinates
// (random.h:57-72); cos_hemisphere() is sample_cos_hemisphere (random.h:80-83) through
// spherical_dir.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace imba {

struct Rng {
    uint64_t s;
    __device__ uint32_t next() {
        const uint32_t c = (uint32_t)(s >> 32);
        const uint32_t x = (uint32_t)(s & 0xFFFFFFFFull);
        s = (uint64_t)x * 4294883355ull + c;
        return x ^ c;
    }
    __device__ float uniform() { return (float)next() / 4294967296.0f; }
};

__device__ __forceinline__ Rng seeded(uint32_t seed_base, int index, int pixel) {
    int seed = (int)seed_base;
    for (int k = 0; k < 4; k++) seed = (int)((uint32_t)(33u * (uint32_t)seed) ^ (uint32_t)index);
    Rng r;
    r.s = (uint64_t)(int64_t)seed;
    const int n = (seed % 5) + 16 + pixel % 5;
    for (int k = 0; k < n; k++) r.next();
    return r;
}

// Path state of the photon-mapping producers (32 bytes; pixel_id first so compaction can test
// it).  The path tracer's state (PtState, 40 bytes) is in pt_shade.h.
struct PathState {
    int pixel_id;
    int sample_id;
    uint64_t rng;
    float tp[3];
    int bounces;
};
static_assert(sizeof(PathState) == 32, "PathState layout");

struct Vec3 {
    float x, y, z;
};

__device__ __forceinline__ Vec3 v3(float x, float y, float z) { return Vec3{x, y, z}; }
__device__ __forceinline__ float dot3(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ Vec3 scale3(Vec3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ Vec3 madd3(Vec3 a, Vec3 b, float s) { return v3(a.x + b.x * s, a.y + b.y * s, a.z + b.z * s); }
__device__ __forceinline__ Vec3 normalize3(Vec3 a) { return scale3(a, 1.0f / sqrtf(dot3(a, a))); }

// Tangent t and binormal b completing the unit normal n to an orthonormal frame.
__device__ __forceinline__ void frame(Vec3 n, Vec3& t, Vec3& b) {
    const bool xbig = fabsf(n.x) > fabsf(n.y);
    const float n0 = xbig ? n.x : n.y;
    const float sig = xbig ? -1.0f : 1.0f;
    const float il = 1.0f / sqrtf(n0 * n0 + n.z * n.z);
    if (xbig) { t.x = n.z * sig * il; t.y = 0.0f; } else { t.y = n.z * sig * il; t.x = 0.0f; }
    t.z = n0 * -1.0f * sig * il;
    b = v3(n.y * t.z - n.z * t.y, n.z * t.x - n.x * t.z, n.x * t.y - n.y * t.x);
    b = scale3(b, 1.0f / sqrtf(dot3(b, b)));
}

// Cosine-weighted direction in local coordinates (z = cos theta).
__device__ __forceinline__ Vec3 cos_hemisphere(float u1, float u2) {
    const float sin_t = sqrtf(1.0f - u2), cos_t = sqrtf(u2);
    float sp, cp;
    sincosf(2.0f * 3.14159265359f * u1, &sp, &cp);
    return v3(sin_t * cp, sin_t * sp, cos_t);
}

}  // namespace imba
