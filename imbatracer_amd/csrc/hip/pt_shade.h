// One path-tracing step of one queue entry: the body of PathTracer::process_primary_rays
// (render/integrators/pt.cpp:82-160) for the materials of the measured scenes — Lambertian
// surfaces (render/materials/brdfs.h:8-20 inside the BSDF of render/materials/bsdf.h:95-231)
// and triangle emitters (render/light.h:13-43, 102-170) — shared by the single-queue and the
// batched launches.  Per hit, in the reference's order and with its arithmetic:
//   - a miss ends the path (no environment map);
//   - an emitter adds throughput * radiance * MIS weight (pt.cpp:120-136) and ends the path;
//   - otherwise next-event estimation (compute_direct_illum, pt.cpp:17-43): one light picked
//     with random_int, a point on it sampled (TriangleLight::sample_direct, light.h:143-170),
//     and a shadow ray written to slot i of the shadow queue with its contribution (added by
//     k_pt_shadow if the ray is unoccluded); a slot without a shadow ray gets pixel_id -1 and a
//     ray that the any-hit traversal ends at the root;
//   - then the bounce (pt.cpp:46-80): path-length cut, Russian roulette (random.h:183-188; the
//     number is drawn on every bounce as pt.cpp:54 does, and the roulette applies from bounce
//     rr_start on, 0 being the reference), BSDF::sample of the Lambertian (component choice and
//     cosine-weighted direction, bsdf.h:147-199, random.h:80-83), throughput update, and the
//     continuation ray from the hit with tmin = 1e-3 t.
// The scenes give every triangle its geometric normal, normalize(cross(v1 - v0, v2 - v0)), as
// its shading normal, and their instances are identities, so calculate_intersection
// (integrator.h:72-122) reduces to isect.normal = isect.geom_normal = normalize(tri normal).
// The device code is compiled without contraction (-ffp-contract=off) and with correctly
// rounded division and square root, so every step but sinf / cosf rounds as the host's does.
#pragma once

#include <hip/hip_runtime.h>

#include "imba_hip.h"
#include "sampling.h"

namespace imba {

constexpr float kPi = 3.14159265359f;  // core/common.h:9
constexpr float kFltMax = 3.402823466e38f;

// Path state of the path tracer (PTState, render/integrators/pt.h: RayState + throughput,
// bounces, last_specular, last_pdf); pixel_id first so compaction can test it.  40 bytes.
struct PtState {
    int pixel_id;
    int sample_id;
    uint64_t rng;
    float tp[3];
    int bounces;
    float last_pdf;
    int last_specular;
};
static_assert(sizeof(PtState) == 40, "PtState layout");

// State of a shadow ray (ShadowState: pixel and contribution); pixel_id = -1 marks an empty
// slot.  16 bytes.
struct ShadowState {
    int pixel_id;
    float tp[3];
};
static_assert(sizeof(ShadowState) == 16, "ShadowState layout");

// local_coordinates (random.h:57-72): tangent t and binormal b of the unit normal n, both
// normalized at the end as the reference does.
__device__ __forceinline__ void local_frame(Vec3 n, Vec3& t, Vec3& b) {
    const bool xbig = fabsf(n.x) > fabsf(n.y);
    const float n0 = xbig ? n.x : n.y;
    const float sig = xbig ? -1.0f : 1.0f;
    const float inv_len = 1.0f / sqrtf(n0 * n0 + n.z * n.z);
    Vec3 tt;
    if (xbig) {
        tt.x = n.z * sig * inv_len;
        tt.y = 0.0f;
    } else {
        tt.y = n.z * sig * inv_len;
        tt.x = 0.0f;
    }
    tt.z = n0 * -1.0f * sig * inv_len;
    Vec3 bb = v3(n.y * tt.z - n.z * tt.y, n.z * tt.x - n.x * tt.z, n.x * tt.y - n.y * tt.x);
    t = scale3(tt, 1.0f / sqrtf(dot3(tt, tt)));
    b = scale3(bb, 1.0f / sqrtf(dot3(bb, bb)));
}

// random_int(min, max) (random.h:31-33) and random_float() (random.h:24-26).
__device__ __forceinline__ int random_int(Rng& r, int lo, int hi) {
    return hi == lo ? lo : (int)(r.next() % (uint32_t)(hi - lo)) + lo;
}

// Checks a shading description (defined in pt.hip).
int check_shading(const ImbaPtShading* sh);

struct PtShadeArgs {
    const float4* rays;
    const int4* hits;
    PtState* states;
    float4* rays_out;        // may alias rays
    float4* shadow_rays;     // slot i: the shadow ray of entry i (or null: no next-event estimation)
    ShadowState* shadow_states;
    float* image;            // 3 floats per pixel, or null
    ImbaPtShading sh;
};

// The Lambertian BSDF at a hit with shading frame (vt, ut, n) = (v_tangent, u_tangent,
// normal) and geometric normal g (= n here): eval and pdf of the pair (out, in)
// (BSDF::eval / BSDF::pdf, bsdf.h:125-145, 202-219; Lambertian::eval, BxDF::pdf).
__device__ __forceinline__ bool lambert_pair(Vec3 n, Vec3 g, Vec3 out, Vec3 in, float& cos_in_local) {
    if (dot3(in, g) * dot3(out, g) <= 0.0f) return false;  // would be the (absent) BTDF
    const float lo = dot3(n, out);
    const float li = dot3(n, in);
    cos_in_local = li;
    return lo * li > 0.0f;  // same_hemisphere in shading space
}

// An empty shadow slot: pixel_id -1 and a ray with tmin > tmax, which the any-hit traversal
// finishes at the root without a hit.
__device__ __forceinline__ void empty_shadow(const PtShadeArgs& a, int i) {
    a.shadow_rays[2 * (size_t)i] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    a.shadow_rays[2 * (size_t)i + 1] = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
    ShadowState e;
    e.pixel_id = -1;
    e.tp[0] = e.tp[1] = e.tp[2] = 0.0f;
    a.shadow_states[i] = e;
}

__device__ __forceinline__ void pt_shade_body(const PtShadeArgs& a, int i, bool& shadow_written);

__device__ __forceinline__ void pt_shade_entry(const PtShadeArgs& a, int i) {
    bool shadow_written = false;
    pt_shade_body(a, i, shadow_written);
    if (a.shadow_rays && !shadow_written) empty_shadow(a, i);
}

__device__ __forceinline__ void pt_shade_body(const PtShadeArgs& a, int i, bool& shadow_written) {
    PtState s = a.states[i];
    const int4 h = a.hits[i];
    const ImbaPtShading& sh = a.sh;
    if (h.x < 0 || h.x >= sh.tri_count) {  // a miss: no environment map, the path ends
        s.pixel_id = -1;
        a.states[i] = s;
        return;
    }
    const int mat = sh.tri_materials[h.x];
    if (mat < 0 || mat >= sh.material_count) {  // malformed scene data: end the path
        s.pixel_id = -1;
        a.states[i] = s;
        return;
    }
    const float4 o = a.rays[2 * (size_t)i];
    const float4 d = a.rays[2 * (size_t)i + 1];
    const float t = __int_as_float(h.z);
    // calculate_intersection: pos = org + tmax * dir, out_dir = -normalize(dir)
    const Vec3 pos = v3(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z);
    const Vec3 dv = v3(d.x, d.y, d.z);
    const Vec3 dn = scale3(dv, 1.0f / sqrtf(dot3(dv, dv)));
    const Vec3 out = v3(-dn.x, -dn.y, -dn.z);
    const Vec3 nraw = v3(sh.tri_normals[3 * (size_t)h.x], sh.tri_normals[3 * (size_t)h.x + 1],
                         sh.tri_normals[3 * (size_t)h.x + 2]);
    const Vec3 n = scale3(nraw, 1.0f / sqrtf(dot3(nraw, nraw)));
    const Vec3 g = n;

    const float er = sh.emission[3 * mat], eg = sh.emission[3 * mat + 1], eb = sh.emission[3 * mat + 2];
    if (er != 0.0f || eg != 0.0f || eb != 0.0f) {
        // AreaEmitter::radiance (light.h:19-39) and the MIS weight of pt.cpp:124-133
        const float cos_o = dot3(g, out);
        float li_r = 0.0f, li_g = 0.0f, li_b = 0.0f, pdf_direct_a = 1.0f;
        if (cos_o > 0.0f) {
            pdf_direct_a = 1.0f / sh.emitter_area[mat];
            li_r = er;
            li_g = eg;
            li_b = eb;
        }
        float pdf_di = pdf_direct_a / (float)sh.light_count;
        const float d_sqr = t * t;
        const float cos_light = dot3(n, out);
        pdf_di *= d_sqr / cos_light;
        const float mis = (s.bounces == 0 || s.last_specular) ? 1.0f : s.last_pdf / (s.last_pdf + pdf_di);
        if (a.image) {
            float* px = a.image + 3 * (size_t)s.pixel_id;
            atomicAdd(px + 0, s.tp[0] * li_r * mis);
            atomicAdd(px + 1, s.tp[1] * li_g * mis);
            atomicAdd(px + 2, s.tp[2] * li_b * mis);
        }
        s.pixel_id = -1;
        a.states[i] = s;
        return;
    }

    Vec3 ut, vt;  // calculate_intersection: local_coordinates(normal, u_tangent, v_tangent)
    local_frame(n, ut, vt);
    const float alb_r = sh.albedo[3 * mat] * (1.0f / kPi);
    const float alb_g = sh.albedo[3 * mat + 1] * (1.0f / kPi);
    const float alb_b = sh.albedo[3 * mat + 2] * (1.0f / kPi);
    Rng rng;
    rng.s = s.rng;

    // compute_direct_illum (pt.cpp:17-43)
    if (a.shadow_rays && sh.light_count > 0) {
        const int l = random_int(rng, 0, sh.light_count);
        const float* L = sh.lights + 12 * (size_t)l;
        const float pdf_lightpick = 1.0f / (float)sh.light_count;
        // TriangleLight: normal_ = cross(p1 - p0, p2 - p0), area = |normal_| / 2, then normalized
        const Vec3 p0 = v3(L[0], L[1], L[2]), p1 = v3(L[3], L[4], L[5]), p2 = v3(L[6], L[7], L[8]);
        const Vec3 e1 = v3(p1.x - p0.x, p1.y - p0.y, p1.z - p0.z), e2 = v3(p2.x - p0.x, p2.y - p0.y, p2.z - p0.z);
        const Vec3 ln = v3(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
        const float area = sqrtf(dot3(ln, ln)) * 0.5f;
        const Vec3 lnn = scale3(ln, 1.0f / sqrtf(dot3(ln, ln)));
        // sample_direct: a uniform point on the triangle (sample_uniform_triangle, random.h:115-119)
        const float r1 = rng.uniform(), r2 = rng.uniform();
        const float sr1 = sqrtf(r1);
        const float lu = 1.0f - sr1, lv = r2 * sr1, lw = 1.0f - lu - lv;
        const Vec3 lp = v3(lu * p0.x + lv * p1.x + lw * p2.x, lu * p0.y + lv * p1.y + lw * p2.y,
                           lu * p0.z + lv * p1.z + lw * p2.z);
        Vec3 dir = v3(lp.x - pos.x, lp.y - pos.y, lp.z - pos.z);
        const float distsq = dot3(dir, dir);
        const float dist = sqrtf(distsq);
        dir = scale3(dir, 1.0f / dist);
        const float cos_out = dot3(lnn, v3(-1.0f * dir.x, -1.0f * dir.y, -1.0f * dir.z));
        float rad_r = 0.0f, rad_g = 0.0f, rad_b = 0.0f, pdf_direct_w = 1.0f;
        if (cos_out > 0.0f && cos_out < 1.0f) {
            const float k = area / distsq;
            rad_r = L[9] * cos_out * k;
            rad_g = L[10] * cos_out * k;
            rad_b = L[11] * cos_out * k;
            pdf_direct_w = 1.0f / area * distsq / cos_out;
        }
        float cos_in = 0.0f;
        const bool refl = lambert_pair(n, g, out, dir, cos_in);
        const float br = refl ? alb_r : 0.0f, bg = refl ? alb_g : 0.0f, bb = refl ? alb_b : 0.0f;
        const float pdf_hit = refl ? fabsf(cos_in) / kPi : 0.0f;
        const float pdf_di = pdf_lightpick * pdf_direct_w;
        const float mis = pdf_di / (pdf_di + pdf_hit);
        if (pdf_hit != 0.0f && pdf_di != 0.0f) {
            const float c = fabsf(dot3(n, dir));
            ShadowState ss;
            ss.pixel_id = s.pixel_id;
            ss.tp[0] = s.tp[0] * br * c * rad_r * mis / pdf_lightpick;
            ss.tp[1] = s.tp[1] * bg * c * rad_g * mis / pdf_lightpick;
            ss.tp[2] = s.tp[2] * bb * c * rad_b * mis / pdf_lightpick;
            const float off = 1e-3f * (dist == kFltMax ? 1.0f : dist);
            a.shadow_rays[2 * (size_t)i] = make_float4(pos.x, pos.y, pos.z, off);
            a.shadow_rays[2 * (size_t)i + 1] = make_float4(dir.x, dir.y, dir.z, dist - off);
            a.shadow_states[i] = ss;
            shadow_written = true;
        }
    }

    // bounce (pt.cpp:46-80)
    bool alive = s.bounces + 1 < sh.max_path_len;
    float rr_pdf = 1.0f;
    if (alive) {
        const float u = rng.uniform();
        if (s.bounces >= sh.rr_start) {
            const float lum = s.tp[0] * 0.2126f + s.tp[1] * 0.7152f + s.tp[2] * 0.0722f;
            rr_pdf = fminf(1.0f, lum * 2.0f);
            alive = u < rr_pdf;
        }
    }
    float pdf = 0.0f;
    Vec3 in = v3(0.0f, 0.0f, 0.0f);
    if (alive) {
        // BSDF::sample: local out direction, component choice (the BRDF: btdf_prob = 0), then
        // BxDF::sample's cosine-weighted direction flipped to the out direction's side.  The
        // reference draws the two numbers of sample_cos_hemisphere (and of
        // sample_uniform_triangle above) as two arguments of one call, whose evaluation order
        // C++ leaves open; they are taken left to right here (as clang compiles that call).
        const float lo_z = dot3(n, out);
        (void)rng.uniform();
        const float u1 = rng.uniform(), u2 = rng.uniform();
        const float sin_t = sqrtf(1.0f - u2), cos_t = sqrtf(u2);
        const float phi = 2.0f * kPi * u1;
        float lx = sin_t * cosf(phi), ly = sin_t * sinf(phi), lz = cos_t;
        pdf = lz / kPi;
        if (lo_z < 0.0f) lz = -lz;
        const bool same = lo_z * lz > 0.0f;
        // local_to_world (bsdf.h:227-231): x along v_tangent, y along u_tangent, z along n
        in = v3(vt.x * lx + ut.x * ly + n.x * lz, vt.y * lx + ut.y * ly + n.y * lz, vt.z * lx + ut.z * ly + n.z * lz);
        if (pdf == 0.0f || dot3(in, g) * dot3(out, g) <= 0.0f || !same) {
            alive = false;  // sampled_flags == BSDF_NONE, or a black value (not the same hemisphere)
        } else if (alb_r <= 0.0f && alb_g <= 0.0f && alb_b <= 0.0f) {
            alive = false;  // is_black(bsdf_value) (core/rgb.h:12-14)
        }
    }
    if (!alive) {
        s.pixel_id = -1;
        s.rng = rng.s;
        a.states[i] = s;
        return;
    }
    const float cos_term = fabsf(dot3(n, in));
    const float q = pdf * rr_pdf;
    s.tp[0] = s.tp[0] * (alb_r * cos_term / q);
    s.tp[1] = s.tp[1] * (alb_g * cos_term / q);
    s.tp[2] = s.tp[2] * (alb_b * cos_term / q);
    s.bounces++;
    s.last_specular = 0;
    s.last_pdf = pdf;
    s.rng = rng.s;
    a.states[i] = s;
    const float off = 1e-3f * t;
    a.rays_out[2 * (size_t)i] = make_float4(pos.x, pos.y, pos.z, off);
    a.rays_out[2 * (size_t)i + 1] = make_float4(in.x, in.y, in.z, kFltMax);
}

// process_shadow_rays (integrator.h:51-66): an unoccluded shadow ray adds its contribution.
__device__ __forceinline__ void pt_shadow_entry(const int4* hits, const ShadowState* st, float* image, int i) {
    const ShadowState s = st[i];
    if (s.pixel_id < 0 || hits[i].x >= 0) return;
    float* px = image + 3 * (size_t)s.pixel_id;
    atomicAdd(px + 0, s.tp[0]);
    atomicAdd(px + 1, s.tp[1]);
    atomicAdd(px + 2, s.tp[2]);
}

}  // namespace imba
