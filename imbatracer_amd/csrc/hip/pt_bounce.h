// One Lambertian path-tracing bounce of one queue entry, shared by the single-queue and the
// batched launches.  PathTracer::bounce (render/integrators/pt.cpp:46-80) for a Lambertian
// material: terminate past the maximum path length, Russian roulette (render/random.h:183-188)
// from bounce `rr_start` on, cosine-weighted direction (random.h:80-83) in the frame of
// random.h:57-72 around the geometric normal facing the ray, throughput times the albedo.  A
// ray that hits an emissive material adds throughput * emission to the image and ends; ended
// paths get pixel_id = -1 (integrator.h:125-128).
#pragma once

#include <hip/hip_runtime.h>

#include "sampling.h"

namespace imba {

struct BounceArgs {
    const float4* rays;
    const int4* hits;
    PtState* states;
    float4* rays_out;  // may alias rays
    int count;         // entries (or, with count_dev, the capacity)
    const int* count_dev;  // device entry count read when the launch runs, or null
    const float* tri_normals;  // 3 per global triangle
    const int* tri_material;
    int n_tris;
    const float* albedo;    // 3 per material
    const float* emission;  // 3 per material
    int n_materials;
    int max_path_len;
    int rr_start;
    float offset;
    float* image;  // 3 per pixel or null
};

__device__ __forceinline__ void bounce_entry(const BounceArgs& a, int i) {
    PtState s = a.states[i];
    const int4 h = a.hits[i];
    if (h.x < 0 || h.x >= a.n_tris) {
        s.pixel_id = -1;
        a.states[i] = s;
        return;
    }
    const int mat = a.tri_material[h.x];
    const bool mat_ok = mat >= 0 && mat < a.n_materials;
    if (mat_ok) {
        const float er = a.emission[3 * mat], eg = a.emission[3 * mat + 1], eb = a.emission[3 * mat + 2];
        if (er > 0.0f || eg > 0.0f || eb > 0.0f) {
            if (a.image) {
                atomicAdd(&a.image[3 * (size_t)s.pixel_id + 0], s.tp[0] * er);
                atomicAdd(&a.image[3 * (size_t)s.pixel_id + 1], s.tp[1] * eg);
                atomicAdd(&a.image[3 * (size_t)s.pixel_id + 2], s.tp[2] * eb);
            }
            s.pixel_id = -1;
            a.states[i] = s;
            return;
        }
    }
    if (s.bounces + 1 >= a.max_path_len || !mat_ok) {
        s.pixel_id = -1;
        a.states[i] = s;
        return;
    }
    Rng rng;
    rng.s = s.rng;
    if (s.bounces >= a.rr_start) {
        const float lum = 0.2126f * s.tp[0] + 0.7152f * s.tp[1] + 0.0722f * s.tp[2];
        const float p = fminf(1.0f, lum * 2.0f);
        if (!(rng.uniform() < p)) {
            s.pixel_id = -1;
            a.states[i] = s;
            return;
        }
        const float ip = 1.0f / p;
        s.tp[0] *= ip; s.tp[1] *= ip; s.tp[2] *= ip;
    }
    const float4 o = a.rays[2 * (size_t)i];
    const float4 d = a.rays[2 * (size_t)i + 1];
    const float t = __int_as_float(h.z);
    const float px = o.x + t * d.x, py = o.y + t * d.y, pz = o.z + t * d.z;
    float nx = a.tri_normals[3 * (size_t)h.x], ny = a.tri_normals[3 * (size_t)h.x + 1], nz = a.tri_normals[3 * (size_t)h.x + 2];
    if (nx * d.x + ny * d.y + nz * d.z > 0.0f) { nx = -nx; ny = -ny; nz = -nz; }
    Vec3 tg, bn;
    const Vec3 nv = v3(nx, ny, nz);
    frame(nv, tg, bn);
    const float u1 = rng.uniform(), u2 = rng.uniform();
    const Vec3 l = cos_hemisphere(u1, u2);
    const float wx = tg.x * l.x + bn.x * l.y + nx * l.z;
    const float wy = tg.y * l.x + bn.y * l.y + ny * l.z;
    const float wz = tg.z * l.x + bn.z * l.y + nz * l.z;
    s.tp[0] *= a.albedo[3 * mat];
    s.tp[1] *= a.albedo[3 * mat + 1];
    s.tp[2] *= a.albedo[3 * mat + 2];
    s.bounces++;
    s.rng = rng.s;
    a.states[i] = s;
    a.rays_out[2 * (size_t)i] = make_float4(px, py, pz, a.offset);
    a.rays_out[2 * (size_t)i + 1] = make_float4(wx, wy, wz, 3.402823466e38f);
}

}  // namespace imba
