// Vertex connection and merging (VCMIntegrator<algo>, render/integrators/vcm.cpp) on the
// device, between traversals: every step the reference runs on the host around RayQueue for
// one iteration, for the materials of the measured scenes (Lambertian, mirror, glass) and
// triangle lights.  The partial MIS weights dVC / dVCM / dVM follow the VCM technical report
// as vcm.cpp computes them (mis_pow is the identity: power 1), with its float operation order.
//
// Light paths (trace_light_paths :64-96, process_light_rays :196-256, bounce :129-193):
//   start    path i belongs to light l (UniformLightTileGen: contiguous runs of path_count /
//            light_count paths, the remainder to light 0), seeded as LightRayGen
//            (render/ray_gen/ray_gen.h:138-145); TriangleLight::sample_emit
//            (render/light.h:109-141); throughput = radiance / pdf_lightpick, path_length 1,
//            dVCM = pdf_direct_a / pdf_emit_w, dVC = cos_out / (pdf_emit_w pdf_lightpick),
//            dVM = dVC eta_vc.
//   at hit   dVCM *= t^2, then dVCM, dVC, dVM *= 1 / |cos|; a vertex on a non-specular
//            material is stored (LightPathVertex: position, direction back along the ray,
//            throughput, dVC, dVCM, dVM, path_length + 1) into slot (bounce, entry) and
//            connected to the camera (connect_to_camera :258-318: one shadow ray to the
//            camera with its MIS-weighted contribution, written to shadow slot i); then the
//            bounce.
// Camera paths (trace_camera_paths :98-127, process_camera_rays :320-443):
//   start    PixelRayGen seeding, jittered pinhole ray; dVCM = light_path_count / pdf_cam_w.
//   at hit   dVCM *= t^2 / |cos|, dVC, dVM *= 1 / |cos|; an emitter adds its MIS-weighted
//            radiance and ends the path; past max_path_len the path ends; otherwise
//            direct_illum (:445-488: one light, one point, one shadow ray, MIS weight),
//            connect (:490-570: num_connections light vertices drawn from the cache, one
//            shadow ray each), a merge query (vertex_merging :572-611, answered by the fused
//            grid query + merge below) and the bounce.
//   bounce   Russian roulette (its number drawn on every bounce, render/random.h:183-188,
//            applied from bounce rr_start on; 0 as the reference), BSDF sample, partial
//            weights (specular: dVCM = 0, dVC and dVM *= cos; otherwise the report's
//            recurrences), throughput *= value cos / (rr_pdf pdf), continuation ray from the
//            hit with tmin = 1e-4 t.
// Merging: the k nearest light vertices within the radius (HashGrid::query) are weighted with
// the Epanechnikov kernel and the MIS weight 1 / (w_light + 1 + w_camera).
// The sub-algorithms switch techniques off as the reference's template parameter does: BPT
// (eta_vm = 0, no merging), LT (no stored vertices, camera connections with weight 1, no camera
// paths) and the VCM integrator's PT (next event with weight 1, no connections or merging, no
// light paths); which passes run is the caller's part.
//
// Materials: kind 0 Lambertian (albedo; render/materials/brdfs.h:8-20 inside the BSDF of
// render/materials/bsdf.h:117-231), 1 mirror (perfect reflection with reflectance albedo),
// 2 glass (index of refraction 1.5, the Fresnel reflectance chooses reflection or refraction,
// throughput unchanged).  Mirror and glass are specular: no vertex is stored, connected or
// merged there, and direct_illum draws its numbers but finds a zero BSDF (no shadow ray).
// Geometric normals are the triangles' own and serve as shading normals; the instances of the
// measured scenes are identities.
#include <hip/hip_runtime.h>

#include "common.h"
#include "grid_query.h"
#include "pt_shade.h"
#include "sampling.h"

namespace imba {
namespace {

constexpr float kGlassEta = 1.5f;
enum { kDiffuse = 0, kMirror = 1, kGlass = 2 };
enum { kAlgoVcm = IMBA_VCM_ALGO_VCM, kAlgoBpt = IMBA_VCM_ALGO_BPT, kAlgoLt = IMBA_VCM_ALGO_LT, kAlgoPt = IMBA_VCM_ALGO_PT };

// VCMState (vcm.h:22-33): RayState + throughput, path_length, finite_light (always true for
// triangle lights), dVC, dVCM, dVM.  48 bytes; pixel_id (light paths: the path index) first,
// -1 once the path ended.
struct VcmState {
    int pixel_id;
    int sample_id;
    uint64_t rng;
    float tp[3];
    int path_length;
    float dVC, dVCM, dVM;
    int pad;
};
static_assert(sizeof(VcmState) == IMBA_VCM_STATE_BYTES, "VcmState layout");

// Merge record of a light vertex in grid slot order: direction back along the light ray,
// dVCM | throughput, dVM.
struct MergeRec {
    float4 out_dvcm;
    float4 tp_dvm;
};

struct Scene {
    const float* tri_normals;
    const int* tri_materials;
    int tri_count;
    const float* albedo;
    const float* emission;
    const float* emitter_area;
    const int* kinds;
    int material_count;
    const float* lights;
    int light_count;
};

struct Frame {
    Scene sc;
    float r2w[16];  // raster -> world (camera.h:35-51), applied to (raster y, raster x, 0, 1)
    float w2r[16];  // world -> raster
    float cam_pos[3];
    float cam_dir[3];
    float img_plane_dist;
    int width, height, spp;
    int light_paths;
    float eta_vc, eta_vm;
    float base_radius;
    float radius_sqr;
    int max_path_len;
    int light_max_depth;
    int rr_start;
    int num_connections;
    int algo;
};

__device__ __forceinline__ Vec3 sub3(Vec3 a, Vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ Vec3 neg3(Vec3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ Vec3 div3(Vec3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }

__device__ __forceinline__ int entries(int capacity, const int* count_dev) {
    return count_dev ? min(max(*count_dev, 0), capacity) : capacity;
}

__device__ __forceinline__ Vec3 tri_normal(const Scene& sc, int tri) {
    const Vec3 nr = v3(sc.tri_normals[3 * (size_t)tri], sc.tri_normals[3 * (size_t)tri + 1],
                       sc.tri_normals[3 * (size_t)tri + 2]);
    return scale3(nr, 1.0f / sqrtf(dot3(nr, nr)));  // normalize (core/float3.h:126-128)
}

// Lambertian BSDF::eval / BSDF::pdf pair at a vertex with shading normal n = geometric normal:
// the reflection lobe applies when both directions lie on one side of the surface.
// pdf(out, in) = |cos in| / pi (BxDF::pdf with cos_hemisphere_pdf).
__device__ __forceinline__ bool lambert_ok(Vec3 n, Vec3 a, Vec3 b) {
    return dot3(b, n) * dot3(a, n) > 0.0f && dot3(n, a) * dot3(n, b) > 0.0f;
}

// Light of path i: contiguous runs of light_paths / light_count, the remainder to light 0.
__device__ __forceinline__ int light_of_path(int i, int paths, int lights) {
    const int per = paths / lights;
    const int first = per + paths % lights;
    if (i < first) return 0;
    return min(lights - 1, 1 + (i - first) / max(per, 1));
}

__global__ __launch_bounds__(256) void k_vcm_light(Frame f, uint32_t seed_base, int first, int count, float4* rays,
                                                   VcmState* states) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= count) return;
    const int i = first + j;
    const Scene& sc = f.sc;
    const int l = light_of_path(i, f.light_paths, sc.light_count);
    const float* L = sc.lights + 12 * (size_t)l;
    Rng rng = seeded(seed_base, i, i);
    const Vec3 p0 = v3(L[0], L[1], L[2]), p1 = v3(L[3], L[4], L[5]), p2 = v3(L[6], L[7], L[8]);
    // TriangleLight: normal_ = cross(p1 - p0, p2 - p0), area = |normal_| / 2, normalized
    const Vec3 e1 = sub3(p1, p0), e2 = sub3(p2, p0);
    const Vec3 ln = v3(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
    const float area = sqrtf(dot3(ln, ln)) * 0.5f;
    const Vec3 n = scale3(ln, 1.0f / sqrtf(dot3(ln, ln)));
    Vec3 tg, bn;
    local_frame(n, tg, bn);
    const float r1 = rng.uniform(), r2 = rng.uniform();
    const float sr1 = sqrtf(r1);
    const float u = 1.0f - sr1, v = r2 * sr1, w = 1.0f - u - v;
    const Vec3 pos = v3(u * p0.x + v * p1.x + w * p2.x, u * p0.y + v * p1.y + w * p2.y, u * p0.z + v * p1.z + w * p2.z);
    const float u1 = rng.uniform(), u2 = rng.uniform();
    const float sin_t = sqrtf(1.0f - u2), cos_t = sqrtf(u2);
    const float phi = 2.0f * kPi * u1;
    const Vec3 ld = v3(sin_t * cosf(phi), sin_t * sinf(phi), cos_t);
    const float dpdf = ld.z / kPi;
    const Vec3 d = v3(ld.x * bn.x + ld.y * tg.x + ld.z * n.x, ld.x * bn.y + ld.y * tg.y + ld.z * n.y,
                      ld.x * bn.z + ld.y * tg.z + ld.z * n.z);
    float rad[3], cos_out, pdf_emit_w, pdf_direct_a;
    if (dpdf <= 0.0f) {
        rad[0] = rad[1] = rad[2] = 0.0f;
        cos_out = 0.0f;
        pdf_emit_w = 1.0f;
        pdf_direct_a = 1.0f;
    } else {
        for (int c = 0; c < 3; c++) rad[c] = L[9 + c] * area * kPi;
        cos_out = ld.z;
        pdf_emit_w = dpdf / area;
        pdf_direct_a = 1.0f / area;
    }
    const float pdf_lightpick = 1.0f / (float)sc.light_count;
    rays[2 * (size_t)j] = make_float4(pos.x, pos.y, pos.z, 1e-3f);
    rays[2 * (size_t)j + 1] = make_float4(d.x, d.y, d.z, kFltMax);
    VcmState s;
    s.pixel_id = i;
    s.sample_id = 0;
    s.rng = rng.s;
    for (int c = 0; c < 3; c++) s.tp[c] = rad[c] / pdf_lightpick;
    s.path_length = 1;
    s.dVCM = pdf_direct_a / pdf_emit_w;
    s.dVC = cos_out / (pdf_emit_w * pdf_lightpick);
    s.dVM = s.dVC * f.eta_vc;
    s.pad = 0;
    states[j] = s;
}

__global__ __launch_bounds__(256) void k_vcm_camera(Frame f, uint32_t seed_base, int first, int count, float4* rays,
                                                    VcmState* states) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= count) return;
    const int i = first + j;
    const int pixel = i / f.spp;
    const int y = pixel / f.width, x = pixel % f.width;
    Rng rng = seeded(seed_base, i, pixel);
    const float sx = (float)x + rng.uniform();
    const float sy = (float)y + rng.uniform();
    // PerspectiveCamera::raster_to_world applied to (raster y, raster x, 0, 1) (camera.h:71-74)
    float t[4];
    const float* m = f.r2w;
    for (int r = 0; r < 4; r++) t[r] = m[4 * r] * sy + m[4 * r + 1] * sx + m[4 * r + 2] * 0.0f + m[4 * r + 3];
    Vec3 d = v3(t[0] / t[3] - f.cam_pos[0], t[1] / t[3] - f.cam_pos[1], t[2] / t[3] - f.cam_pos[2]);
    d = scale3(d, 1.0f / sqrtf(dot3(d, d)));
    rays[2 * (size_t)j] = make_float4(f.cam_pos[0], f.cam_pos[1], f.cam_pos[2], 0.0f);
    rays[2 * (size_t)j + 1] = make_float4(d.x, d.y, d.z, kFltMax);
    const float cos_o = dot3(d, v3(f.cam_dir[0], f.cam_dir[1], f.cam_dir[2]));
    const float r = f.img_plane_dist / cos_o;
    const float pdf_cam_w = r * r / cos_o;
    VcmState s;
    s.pixel_id = pixel;
    s.sample_id = i % f.spp;
    s.rng = rng.s;
    s.tp[0] = s.tp[1] = s.tp[2] = 1.0f;
    s.path_length = 1;
    s.dVC = 0.0f;
    s.dVM = 0.0f;
    s.dVCM = (float)f.light_paths / pdf_cam_w;
    s.pad = 0;
    states[j] = s;
}

// Mirror / glass continuation (as the photon-mapping producers): nf faces the incoming ray's
// side; `entering` is true when the ray arrives on the normal's side.
__device__ __forceinline__ Vec3 specular_dir(int kind, Vec3 d, Vec3 nf, bool entering, Rng& rng) {
    const float cos_i = -dot3(d, nf);
    const Vec3 refl = madd3(d, nf, 2.0f * cos_i);
    if (kind == kMirror) return refl;
    const float eta = entering ? 1.0f / kGlassEta : kGlassEta;
    const float sin2_t = eta * eta * (1.0f - cos_i * cos_i);
    if (sin2_t >= 1.0f) return refl;  // total internal reflection
    const float cos_t = sqrtf(1.0f - sin2_t);
    const float rs = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    const float rp = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    const float fr = 0.5f * (rs * rs + rp * rp);
    if (rng.uniform() < fr) return refl;
    return madd3(scale3(d, eta), nf, eta * cos_i - cos_t);
}

// VCMIntegrator::bounce (vcm.cpp:129-193) at a hit (pos, out = -normalize(dir), n) of
// material mat / kind.  Returns false when the path ends; otherwise the state and the
// continuation direction *in are updated.
// `adjoint` (light paths) takes the cosine through shading_normal_adjoint (vcm.h:115-117).
__device__ __forceinline__ float cos_in(bool adjoint, Vec3 n, Vec3 out, Vec3 in) {
    return adjoint ? fabsf(dot3(out, n) * dot3(in, n) / dot3(out, n)) : fabsf(dot3(in, n));
}

__device__ __forceinline__ bool vcm_bounce(const Frame& f, VcmState& s, Rng& rng, int mat, int kind, Vec3 d, Vec3 out,
                                           Vec3 n, bool adjoint, Vec3& in) {
    const Scene& sc = f.sc;
    float rr_pdf = 1.0f;
    {
        const float u = rng.uniform();
        const float lum = s.tp[0] * 0.2126f + s.tp[1] * 0.7152f + s.tp[2] * 0.0722f;
        const float p = fminf(1.0f, lum * 2.0f);
        if (s.path_length - 1 >= f.rr_start) {
            rr_pdf = p;
            if (!(u < rr_pdf)) return false;
        }
    }
    const float ar = sc.albedo[3 * mat], ag = sc.albedo[3 * mat + 1], ab = sc.albedo[3 * mat + 2];
    if (kind == kDiffuse) {
        // BSDF::sample: component draw (the BRDF only), BxDF::sample's cosine-weighted direction
        // flipped to the out side, local_to_world over (v_tangent, u_tangent, normal)
        Vec3 ut, vt;
        local_frame(n, ut, vt);
        const float lo_z = dot3(n, out);
        (void)rng.uniform();
        const float u1 = rng.uniform(), u2 = rng.uniform();
        const float sin_t = sqrtf(1.0f - u2), cos_t = sqrtf(u2);
        const float phi = 2.0f * kPi * u1;
        float lx = sin_t * cosf(phi), ly = sin_t * sinf(phi), lz = cos_t;
        const float pdf_dir_w = lz / kPi;
        if (lo_z < 0.0f) lz = -lz;
        in = v3(vt.x * lx + ut.x * ly + n.x * lz, vt.y * lx + ut.y * ly + n.y * lz, vt.z * lx + ut.z * ly + n.z * lz);
        if (pdf_dir_w == 0.0f || !lambert_ok(n, out, in)) return false;
        const float vr = ar * (1.0f / kPi), vg = ag * (1.0f / kPi), vb = ab * (1.0f / kPi);
        if (vr <= 0.0f && vg <= 0.0f && vb <= 0.0f) return false;
        const float pdf_rev_w = fabsf(dot3(n, out)) / kPi;  // bsdf->pdf(sample_dir, out_dir)
        const float cos_i = cos_in(adjoint, n, out, in);
        const float c = cos_i / pdf_dir_w;
        s.dVC = c * (s.dVC * pdf_rev_w + s.dVCM + f.eta_vm);
        s.dVM = c * (s.dVM * pdf_rev_w + s.dVCM * f.eta_vc + 1.0f);
        s.dVCM = 1.0f / pdf_dir_w;
        const float q = rr_pdf * pdf_dir_w;
        s.tp[0] = s.tp[0] * ((vr * cos_i) / q);
        s.tp[1] = s.tp[1] * ((vg * cos_i) / q);
        s.tp[2] = s.tp[2] * ((vb * cos_i) / q);
    } else {
        const float cos_o = dot3(d, n);
        const bool entering = cos_o < 0.0f;
        const Vec3 nf = entering ? n : neg3(n);
        in = specular_dir(kind, d, nf, entering, rng);
        const float cos_i = cos_in(adjoint, n, out, in);
        s.dVCM = 0.0f;
        s.dVC *= cos_i;
        s.dVM *= cos_i;
        if (kind == kMirror) {
            s.tp[0] = s.tp[0] * ar / rr_pdf;
            s.tp[1] = s.tp[1] * ag / rr_pdf;
            s.tp[2] = s.tp[2] * ab / rr_pdf;
        } else {
            s.tp[0] = s.tp[0] / rr_pdf;
            s.tp[1] = s.tp[1] / rr_pdf;
            s.tp[2] = s.tp[2] / rr_pdf;
        }
    }
    s.path_length++;
    return true;
}

__device__ __forceinline__ void empty_shadow(float4* rays, ShadowState* st, size_t slot) {
    rays[2 * slot] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    rays[2 * slot + 1] = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
    ShadowState e;
    e.pixel_id = -1;
    e.tp[0] = e.tp[1] = e.tp[2] = 0.0f;
    st[slot] = e;
}

struct LightArgs {
    const float4* rays;
    const int4* hits;
    VcmState* states;
    float4* rays_out;
    int capacity;
    const int* count_dev;
    float4* vtx_a;          // slots of this bounce: position | path_length, out | tri
    float4* vtx_b;          // key | throughput, dVCM dVM dVC material
    float4* shadow_rays;    // camera connections, slot i
    ShadowState* shadow_states;
};

__global__ __launch_bounds__(256) void k_vcm_light_bounce(Frame f, LightArgs a) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.capacity) return;
    // every slot of the capacity is written: an empty vertex slot has key -1
    a.vtx_b[2 * (size_t)i] = make_float4(__int_as_float(-1), 0.0f, 0.0f, 0.0f);
    const int n_entries = entries(a.capacity, a.count_dev);
    bool shadow = false;
    if (i < n_entries) {
        const Scene& sc = f.sc;
        VcmState s = a.states[i];
        const int4 h = a.hits[i];
        bool alive = h.x >= 0 && h.x < sc.tri_count;
        const int mat = alive ? sc.tri_materials[h.x] : -1;
        alive = alive && mat >= 0 && mat < sc.material_count;
        if (alive) {
            const float4 o = a.rays[2 * (size_t)i];
            const float4 dr = a.rays[2 * (size_t)i + 1];
            const float t = __int_as_float(h.z);
            const Vec3 d = v3(dr.x, dr.y, dr.z);
            const Vec3 pos = v3(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z);
            const Vec3 out = neg3(scale3(d, 1.0f / sqrtf(dot3(d, d))));
            const Vec3 n = tri_normal(sc, h.x);
            const float cos_o = fabsf(dot3(out, n));
            if (cos_o == 0.0f) {
                alive = false;
            } else {
                s.dVCM *= t * t;  // path_length > 1 || finite_light: always for triangle lights
                const float ic = 1.0f / cos_o;
                s.dVCM *= ic;
                s.dVC *= ic;
                s.dVM *= ic;
                const int kind = sc.kinds[mat];
                if (kind == kDiffuse) {
                    if (f.algo != kAlgoLt) {
                        a.vtx_a[2 * (size_t)i] = make_float4(pos.x, pos.y, pos.z, __int_as_float(s.path_length + 1));
                        a.vtx_a[2 * (size_t)i + 1] = make_float4(out.x, out.y, out.z, __int_as_float(h.x));
                        a.vtx_b[2 * (size_t)i] = make_float4(__int_as_float(s.pixel_id), s.tp[0], s.tp[1], s.tp[2]);
                        a.vtx_b[2 * (size_t)i + 1] = make_float4(s.dVCM, s.dVM, s.dVC, __int_as_float(mat));
                    }
                    // connect_to_camera (vcm.cpp:258-318)
                    const Vec3 cp = v3(f.cam_pos[0], f.cam_pos[1], f.cam_pos[2]);
                    const Vec3 cd = v3(f.cam_dir[0], f.cam_dir[1], f.cam_dir[2]);
                    Vec3 dtc = sub3(cp, pos);
                    if (dot3(neg3(dtc), cd) >= 0.0f) {
                        const float* m = f.w2r;  // world_to_raster (camera.h:66-69)
                        float r[4];
                        for (int k = 0; k < 4; k++)
                            r[k] = m[4 * k] * pos.x + m[4 * k + 1] * pos.y + m[4 * k + 2] * pos.z + m[4 * k + 3] * 1.0f;
                        const float rx = r[1] / r[3], ry = r[0] / r[3];
                        const float fx = floorf(rx), fy = floorf(ry);
                        const int px = (int)fx, py = (int)fy;
                        if (px >= 0 && px < f.width && py >= 0 && py < f.height) {
                            const int pixel = py * f.width + px;
                            const float d2 = dot3(dtc, dtc);
                            const float dist = sqrtf(d2);
                            dtc = div3(dtc, dist);
                            const float cos_cam = fabsf(dot3(cd, neg3(dtc)));
                            const float cos_surf = fabsf(dot3(out, n) * dot3(dtc, n) / dot3(out, n));
                            if (lambert_ok(n, out, dtc)) {
                                const float pdf_rev_w = fabsf(dot3(n, out)) / kPi;
                                const float ipd = f.img_plane_dist;
                                const float img_to_surf = (ipd * ipd * cos_surf) / (d2 * cos_cam * (cos_cam * cos_cam));
                                const float wl = img_to_surf / (float)f.light_paths *
                                                 (f.eta_vm + s.dVCM + s.dVC * pdf_rev_w);
                                const float mis = f.algo == kAlgoLt ? 1.0f : 1.0f / (wl + 1.0f);
                                const float alb[3] = {sc.albedo[3 * mat] * (1.0f / kPi),
                                                      sc.albedo[3 * mat + 1] * (1.0f / kPi),
                                                      sc.albedo[3 * mat + 2] * (1.0f / kPi)};
                                ShadowState ss;
                                ss.pixel_id = pixel;
                                for (int c = 0; c < 3; c++)
                                    ss.tp[c] = s.tp[c] * (mis * alb[c] * img_to_surf / (float)f.light_paths);
                                const float off = dist * 1e-3f;
                                a.shadow_rays[2 * (size_t)i] = make_float4(pos.x, pos.y, pos.z, off);
                                a.shadow_rays[2 * (size_t)i + 1] = make_float4(dtc.x, dtc.y, dtc.z, dist - off);
                                a.shadow_states[i] = ss;
                                shadow = true;
                            }
                        }
                    }
                }
                if (s.path_length >= f.light_max_depth) {
                    alive = false;
                } else {
                    Rng rng;
                    rng.s = s.rng;
                    Vec3 in;
                    alive = vcm_bounce(f, s, rng, mat, kind, d, out, n, true, in);
                    s.rng = rng.s;
                    if (alive) {
                        a.rays_out[2 * (size_t)i] = make_float4(pos.x, pos.y, pos.z, t * 1e-4f);
                        a.rays_out[2 * (size_t)i + 1] = make_float4(in.x, in.y, in.z, kFltMax);
                    }
                }
            }
        }
        if (!alive) s.pixel_id = -1;
        a.states[i] = s;
    }
    if (!shadow) empty_shadow(a.shadow_rays, a.shadow_states, i);
}

struct CameraArgs {
    const float4* rays;
    const int4* hits;
    VcmState* states;
    float4* rays_out;
    int capacity;
    const int* count_dev;
    float* image;
    // the light vertex cache (dense): count on the device
    const float4* vtx_a;
    const float4* vtx_b;
    const int* vtx_count_dev;
    // shadow slots: (1 + num_connections) per entry, entry i's at (1 + C) i + c
    float4* shadow_rays;
    ShadowState* shadow_states;
    int* shadow_count_dev;  // (1 + C) * entries, written by the launch (nullable)
    // merge queries, slot i: position | pixel, normal | dVCM, out | dVM, throughput | material
    float4* queries;
};

__global__ __launch_bounds__(256) void k_vcm_camera_bounce(Frame f, CameraArgs a) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int n_entries = entries(a.capacity, a.count_dev);
    const int C = f.num_connections;
    if (i == 0 && a.shadow_count_dev) *a.shadow_count_dev = (1 + C) * n_entries;
    if (i >= n_entries) return;
    const Scene& sc = f.sc;
    const size_t sbase = (size_t)(1 + C) * i;
    int written = 0;  // bit c: shadow slot c holds a ray
    bool query = false;
    VcmState s = a.states[i];
    const int4 h = a.hits[i];
    bool alive = h.x >= 0 && h.x < sc.tri_count;
    const int mat = alive ? sc.tri_materials[h.x] : -1;
    alive = alive && mat >= 0 && mat < sc.material_count;
    if (alive) {
        const float4 o = a.rays[2 * (size_t)i];
        const float4 dr = a.rays[2 * (size_t)i + 1];
        const float t = __int_as_float(h.z);
        const Vec3 d = v3(dr.x, dr.y, dr.z);
        const Vec3 pos = v3(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z);
        const Vec3 out = neg3(scale3(d, 1.0f / sqrtf(dot3(d, d))));
        const Vec3 n = tri_normal(sc, h.x);
        const float cos_o = fabsf(dot3(out, n));
        s.dVCM *= t * t / cos_o;
        s.dVC *= 1.0f / cos_o;
        s.dVM *= 1.0f / cos_o;
        const float er = sc.emission[3 * mat], eg = sc.emission[3 * mat + 1], eb = sc.emission[3 * mat + 2];
        const int kind = sc.kinds[mat];
        if (cos_o == 0.0f) {
            alive = false;
        } else if (er != 0.0f || eg != 0.0f || eb != 0.0f) {
            // AreaEmitter::radiance (light.h:19-39) with the geometric normal, MIS of :408-417
            const float c = dot3(n, out);
            float li[3] = {0.0f, 0.0f, 0.0f}, pdf_direct_a = 1.0f, pdf_emit_w = 1.0f;
            if (c > 0.0f) {
                const float ia = 1.0f / sc.emitter_area[mat];
                pdf_direct_a = ia;
                pdf_emit_w = ia * (fabsf(c) / kPi);
                li[0] = er;
                li[1] = eg;
                li[2] = eb;
            }
            const float pick = 1.0f / (float)sc.light_count;
            const float pdf_di = pdf_direct_a * pick, pdf_e = pdf_emit_w * pick;
            const float wc = pdf_di * s.dVCM + pdf_e * s.dVC;
            const float mis = s.path_length == 1 ? 1.0f : 1.0f / (wc + 1.0f);
            if (a.image) {
                float* px = a.image + 3 * (size_t)s.pixel_id;
                atomicAdd(px + 0, s.tp[0] * li[0] * mis);
                atomicAdd(px + 1, s.tp[1] * li[1] * mis);
                atomicAdd(px + 2, s.tp[2] * li[2] * mis);
            }
            alive = false;
        } else if (s.path_length >= f.max_path_len) {
            alive = false;
        } else {
            Rng rng;
            rng.s = s.rng;
            const float alb[3] = {sc.albedo[3 * mat] * (1.0f / kPi), sc.albedo[3 * mat + 1] * (1.0f / kPi),
                                  sc.albedo[3 * mat + 2] * (1.0f / kPi)};
            // direct_illum (vcm.cpp:445-488)
            if (sc.light_count > 0) {
                const int l = random_int(rng, 0, sc.light_count);
                const float* L = sc.lights + 12 * (size_t)l;
                const float pick_inv = (float)sc.light_count;
                const Vec3 p0 = v3(L[0], L[1], L[2]), p1 = v3(L[3], L[4], L[5]), p2 = v3(L[6], L[7], L[8]);
                const Vec3 e1 = sub3(p1, p0), e2 = sub3(p2, p0);
                const Vec3 ln = v3(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
                const float area = sqrtf(dot3(ln, ln)) * 0.5f;
                const Vec3 lnn = scale3(ln, 1.0f / sqrtf(dot3(ln, ln)));
                const float r1 = rng.uniform(), r2 = rng.uniform();
                const float sr1 = sqrtf(r1);
                const float lu = 1.0f - sr1, lv = r2 * sr1, lw = 1.0f - lu - lv;
                const Vec3 lp = v3(lu * p0.x + lv * p1.x + lw * p2.x, lu * p0.y + lv * p1.y + lw * p2.y,
                                   lu * p0.z + lv * p1.z + lw * p2.z);
                Vec3 dir = sub3(lp, pos);
                const float distsq = dot3(dir, dir);
                const float dist = sqrtf(distsq);
                dir = scale3(dir, 1.0f / dist);
                const float lcos = dot3(lnn, v3(-1.0f * dir.x, -1.0f * dir.y, -1.0f * dir.z));
                float rad[3] = {0.0f, 0.0f, 0.0f}, cos_out = 1.0f, pdf_emit_w = 1.0f, pdf_direct_w = 1.0f;
                if (lcos > 0.0f && lcos < 1.0f) {
                    const float k = area / distsq;
                    for (int c = 0; c < 3; c++) rad[c] = L[9 + c] * lcos * k;
                    cos_out = lcos;
                    pdf_emit_w = fabsf(lcos) / kPi / area;
                    pdf_direct_w = 1.0f / area * distsq / lcos;
                }
                if (kind == kDiffuse && lambert_ok(n, out, dir)) {
                    const float cos_i = fabsf(dot3(n, dir));
                    const float pdf_dir_w = fabsf(dot3(n, dir)) / kPi;
                    const float pdf_rev_w = fabsf(dot3(n, out)) / kPi;
                    if (pdf_dir_w != 0.0f && pdf_rev_w != 0.0f) {
                        const float wl = pdf_dir_w * pick_inv / pdf_direct_w;
                        const float wc = pdf_emit_w * cos_i / (pdf_direct_w * cos_out) *
                                         (f.eta_vm + s.dVCM + s.dVC * pdf_rev_w);
                        const float mis = f.algo == kAlgoPt ? 1.0f : 1.0f / (wc + 1.0f + wl);
                        ShadowState ss;
                        ss.pixel_id = s.pixel_id;
                        for (int c = 0; c < 3; c++) ss.tp[c] = s.tp[c] * mis * alb[c] * cos_i * rad[c] * pick_inv;
                        const float off = 1e-3f * (dist == kFltMax ? 1.0f : dist);
                        a.shadow_rays[2 * sbase] = make_float4(pos.x, pos.y, pos.z, off);
                        a.shadow_rays[2 * sbase + 1] = make_float4(dir.x, dir.y, dir.z, dist - off);
                        a.shadow_states[sbase] = ss;
                        written |= 1;
                    }
                }
            }
            if (kind == kDiffuse && (f.algo == kAlgoVcm || f.algo == kAlgoBpt)) {
                // connect (vcm.cpp:490-570)
                const int nv = a.vtx_count_dev ? *a.vtx_count_dev : 0;
                const float vc_weight = (float)nv / ((float)f.light_paths * (float)C);
                for (int c = 0; c < C; c++) {
                    if (nv <= 0) break;
                    const int vi = random_int(rng, 0, nv);
                    const float4 va0 = a.vtx_a[2 * (size_t)vi], va1 = a.vtx_a[2 * (size_t)vi + 1];
                    const float4 vb0 = a.vtx_b[2 * (size_t)vi], vb1 = a.vtx_b[2 * (size_t)vi + 1];
                    const int lpl = __float_as_int(va0.w);
                    if (lpl + s.path_length > f.max_path_len) continue;
                    const Vec3 lpos = v3(va0.x, va0.y, va0.z);
                    const Vec3 lout = v3(va1.x, va1.y, va1.z);
                    const int ltri = __float_as_int(va1.w);
                    const int lmat = __float_as_int(vb1.w);
                    const Vec3 lnrm = tri_normal(sc, ltri);
                    Vec3 cdir = sub3(lpos, pos);
                    const float d2 = dot3(cdir, cdir);
                    const float dist = sqrtf(d2);
                    cdir = scale3(cdir, 1.0f / dist);
                    if (dist < f.base_radius) continue;
                    const Vec3 ncdir = neg3(cdir);
                    if (!lambert_ok(n, out, cdir) || !lambert_ok(lnrm, lout, ncdir)) continue;
                    const float pdf_dir_cam_w = fabsf(dot3(n, cdir)) / kPi;
                    const float pdf_rev_cam_w = fabsf(dot3(n, out)) / kPi;
                    const float pdf_dir_light_w = fabsf(dot3(lnrm, ncdir)) / kPi;
                    const float pdf_rev_light_w = fabsf(dot3(lnrm, lout)) / kPi;
                    if (pdf_dir_cam_w == 0.0f || pdf_dir_light_w == 0.0f || pdf_rev_cam_w == 0.0f ||
                        pdf_rev_light_w == 0.0f)
                        continue;
                    const float cos_cam = fabsf(dot3(n, cdir));
                    const float cos_light = fabsf(dot3(lout, lnrm) * dot3(ncdir, lnrm) / dot3(lout, lnrm));
                    const float geom = cos_cam * cos_light / d2;
                    if (geom <= 0.0f) continue;
                    const float pdf_cam_a = pdf_dir_cam_w * cos_light / d2;
                    const float pdf_light_a = pdf_dir_light_w * cos_cam / d2;
                    const float l_dvcm = vb1.x, l_dvc = vb1.z;
                    const float wl = pdf_cam_a * (f.eta_vm + l_dvcm + l_dvc * pdf_rev_light_w);
                    const float wc = pdf_light_a * (f.eta_vm + s.dVCM + s.dVC * pdf_rev_cam_w);
                    const float mis = 1.0f / (wc + 1.0f + wl);
                    const float lalb[3] = {sc.albedo[3 * lmat] * (1.0f / kPi), sc.albedo[3 * lmat + 1] * (1.0f / kPi),
                                           sc.albedo[3 * lmat + 2] * (1.0f / kPi)};
                    const float ltp[3] = {vb0.y, vb0.z, vb0.w};
                    ShadowState ss;
                    ss.pixel_id = s.pixel_id;
                    for (int k = 0; k < 3; k++)
                        ss.tp[k] = s.tp[k] * vc_weight * mis * geom * alb[k] * lalb[k] * ltp[k];
                    const float off = 1e-3f * dist;
                    const size_t slot = sbase + 1 + c;
                    a.shadow_rays[2 * slot] = make_float4(pos.x, pos.y, pos.z, off);
                    a.shadow_rays[2 * slot + 1] = make_float4(cdir.x, cdir.y, cdir.z, dist - off);
                    a.shadow_states[slot] = ss;
                    written |= 2 << c;
                }
            }
            if (kind == kDiffuse && f.algo == kAlgoVcm) {
                // vertex_merging (vcm.cpp:572-611): the query, answered after this launch
                if (a.queries) {
                    a.queries[4 * (size_t)i] = make_float4(pos.x, pos.y, pos.z, __int_as_float(s.pixel_id));
                    a.queries[4 * (size_t)i + 1] = make_float4(n.x, n.y, n.z, s.dVCM);
                    a.queries[4 * (size_t)i + 2] = make_float4(out.x, out.y, out.z, s.dVM);
                    a.queries[4 * (size_t)i + 3] = make_float4(s.tp[0], s.tp[1], s.tp[2], __int_as_float(mat));
                    query = true;
                }
            }
            Vec3 in;
            alive = vcm_bounce(f, s, rng, mat, kind, d, out, n, false, in);
            s.rng = rng.s;
            if (alive) {
                a.rays_out[2 * (size_t)i] = make_float4(pos.x, pos.y, pos.z, t * 1e-4f);
                a.rays_out[2 * (size_t)i + 1] = make_float4(in.x, in.y, in.z, kFltMax);
            }
        }
    }
    if (!alive) s.pixel_id = -1;
    a.states[i] = s;
    for (int c = 0; c <= C; c++)
        if (!(written & (1 << c))) empty_shadow(a.shadow_rays, a.shadow_states, sbase + c);
    if (a.queries && !query)
        a.queries[4 * (size_t)i] = make_float4(__int_as_float(0x7f800000), __int_as_float(0x7f800000),
                                               __int_as_float(0x7f800000), __int_as_float(-1));
}

// Merge records in grid slot order: rec[slot] from the light vertex order[slot].
__global__ __launch_bounds__(256) void k_vcm_merge_records(const float4* vtx_a, const float4* vtx_b, const int* order,
                                                           int n, MergeRec* rec) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    const int v = order[s];
    const float4 a1 = vtx_a[2 * (size_t)v + 1];
    const float4 b0 = vtx_b[2 * (size_t)v], b1 = vtx_b[2 * (size_t)v + 1];
    MergeRec r;
    r.out_dvcm = make_float4(a1.x, a1.y, a1.z, b1.x);
    r.tp_dvm = make_float4(b0.y, b0.z, b0.w, b1.y);
    rec[s] = r;
}

struct MergeArgs {
    const GridParams* gp;
    const int* ends;
    const float4* pos4;
    const uint32_t* box;
    const MergeRec* rec;
    const float4* queries;
    int capacity;
    const int* count_dev;
    int k;
    const float* albedo;
    float* image;
};

// vertex_merging's sum for one merge query: photon j of the list (its squared distance d2 and
// merge record) weighted by the Epanechnikov kernel and the MIS weight; the wave adds the
// lanes' terms in list order, as the reference's loop does.
__device__ __forceinline__ void merge_epilogue(const Frame& f, const MergeArgs& a, int q, int count, float my_dist,
                                               int my_slot, float k_dist) {
    const int lane = (int)(threadIdx.x & 63);
    const float4 q0 = a.queries[4 * (size_t)q];
    const float4 q1 = a.queries[4 * (size_t)q + 1];
    const float4 q2 = a.queries[4 * (size_t)q + 2];
    const float4 q3 = a.queries[4 * (size_t)q + 3];
    const int pixel = __float_as_int(q0.w);
    const Vec3 n = v3(q1.x, q1.y, q1.z), out = v3(q2.x, q2.y, q2.z);
    const int mat = __float_as_int(q3.w);
    const float rs = count == a.k ? k_dist : f.radius_sqr;
    float cr = 0.0f, cg = 0.0f, cb = 0.0f;
    if (lane < count) {
        const MergeRec r = a.rec[my_slot];
        const Vec3 pin = v3(r.out_dvcm.x, r.out_dvcm.y, r.out_dvcm.z);
        if (lambert_ok(n, out, pin)) {
            const float pdf_dir_w = fabsf(dot3(n, pin)) / kPi;
            const float pdf_rev_w = fabsf(dot3(n, out)) / kPi;
            const float vr = a.albedo[3 * mat] * (1.0f / kPi), vg = a.albedo[3 * mat + 1] * (1.0f / kPi),
                        vb = a.albedo[3 * mat + 2] * (1.0f / kPi);
            if (pdf_dir_w != 0.0f && pdf_rev_w != 0.0f && !(vr <= 0.0f && vg <= 0.0f && vb <= 0.0f)) {
                const float wl = r.out_dvcm.w * f.eta_vc + r.tp_dvm.w * pdf_dir_w;
                const float wc = q1.w * f.eta_vc + q2.w * pdf_rev_w;
                const float mis = 1.0f / (wl + 1.0f + wc);
                const float kern = 1.0f - my_dist / rs;
                cr = mis * vr * kern * r.tp_dvm.x;
                cg = mis * vg * kern * r.tp_dvm.y;
                cb = mis * vb * kern * r.tp_dvm.z;
            }
        }
    }
    // in-order sum over the list (lane 0 first), as contrib += ... in the loop; the terms are
    // read lane by lane into scalar registers
    float sr = 0.0f, sg = 0.0f, sb = 0.0f;
    for (int j = 0; j < count; j++) {
        sr += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cr), j));
        sg += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cg), j));
        sb += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cb), j));
    }
    if (lane == 0 && count > 0 && pixel >= 0) {
        const float scale = 2.0f / (kPi * rs * (float)f.light_paths);
        sr *= scale;
        sg *= scale;
        sb *= scale;
        float* px = a.image + 3 * (size_t)pixel;
        atomicAdd(px + 0, q3.x * sr);
        atomicAdd(px + 1, q3.y * sg);
        atomicAdd(px + 2, q3.z * sb);
    }
}

// Grid query and merge in one launch, one query per wave: the k-capped list keeps each
// photon's grid slot, and the merge reads the slot-ordered records next to the positions the
// scan just read.
__global__ __launch_bounds__(256) void k_vcm_merge(Frame f, MergeArgs a) {
    const int q = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (q >= entries(a.capacity, a.count_dev)) return;  // whole waves leave together
    const float4 q0 = a.queries[4 * (size_t)q];
    if (__float_as_int(q0.w) < 0) return;
    float my_dist;
    int my_slot;
    const int count = query_wave_list<true>(*a.gp, a.ends, a.pos4, a.box, q0.x, q0.y, q0.z, a.k, my_dist, my_slot);
    if (count == 0) return;
    const float kd = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(my_dist), a.k - 1));
    merge_epilogue(f, a, q, count, my_dist, my_slot, kd);
}

int check_frame(const ImbaVcmFrame* p, Frame* f) {
    IMBA_CHECK_ARG(p, "frame is NULL");
    IMBA_CHECK_ARG(p->tri_normals && p->tri_materials && p->albedo && p->emission && p->emitter_area && p->kinds,
                   "NULL scene array");
    IMBA_CHECK_ARG(p->lights && p->light_count > 0, "VCM needs at least one triangle light");
    IMBA_CHECK_ARG(p->tri_count >= 0 && p->material_count > 0, "bad scene sizes");
    IMBA_CHECK_ARG(p->width > 0 && p->height > 0 && p->spp > 0 && p->light_paths > 0, "bad image sizes");
    IMBA_CHECK_ARG(p->num_connections >= 1 && p->num_connections <= 8, "num_connections must be in [1, 8]");
    IMBA_CHECK_ARG(p->max_path_len >= 1 && p->light_max_depth >= 1, "path lengths must be >= 1");
    IMBA_CHECK_ARG(p->radius > 0.0f, "radius must be > 0");
    IMBA_CHECK_ARG(p->algo == kAlgoVcm || p->algo == kAlgoBpt || p->algo == kAlgoLt || p->algo == kAlgoPt,
                   "algo must be IMBA_VCM_ALGO_VCM, _BPT, _LT or _PT (PPM runs through imba_ppm_*)");
    Frame& F = *f;
    F.sc = Scene{p->tri_normals, p->tri_materials, p->tri_count, p->albedo, p->emission, p->emitter_area,
                 p->kinds, p->material_count, p->lights, p->light_count};
    for (int k = 0; k < 16; k++) {
        F.r2w[k] = p->raster_to_world[k];
        F.w2r[k] = p->world_to_raster[k];
    }
    for (int k = 0; k < 3; k++) {
        F.cam_pos[k] = p->cam_pos[k];
        F.cam_dir[k] = p->cam_dir[k];
    }
    F.img_plane_dist = p->image_plane_dist;
    F.width = p->width;
    F.height = p->height;
    F.spp = p->spp;
    F.light_paths = p->light_paths;
    // render() (vcm.cpp:24-56) for the first iteration: pm_radius = base_radius
    const float r = p->radius;
    const float kPiHost = 3.14159265359f;
    const float eta_vcm = kPiHost * (r * r) * (float)p->light_paths;
    F.eta_vc = 1.0f / eta_vcm;
    F.eta_vm = p->algo == kAlgoBpt ? 0.0f : eta_vcm;  // mis_eta_vm_ (vcm.cpp:45)
    F.base_radius = r;
    F.radius_sqr = r * r;
    F.max_path_len = p->max_path_len;
    F.light_max_depth = p->light_max_depth;
    F.rr_start = p->rr_start;
    F.num_connections = p->num_connections;
    F.algo = p->algo;
    return IMBA_OK;
}

}  // namespace
}  // namespace imba

using namespace imba;

extern "C" {

int imba_vcm_light(const ImbaVcmFrame* frame, uint32_t seed_base, int first, int count, ImbaRay* rays, void* states,
                   void* stream) {
    Frame f;
    int r = check_frame(frame, &f);
    if (r) return r;
    IMBA_CHECK_ARG(first >= 0 && count >= 0, "bad sizes");
    if (count == 0) return IMBA_OK;
    IMBA_CHECK_ARG(rays && states, "NULL buffer");
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    hipLaunchKernelGGL(k_vcm_light, dim3(ceil_div(count, 256)), dim3(256), 0, s, f, seed_base, first, count,
                       (float4*)rays, (VcmState*)states);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_vcm_camera(const ImbaVcmFrame* frame, uint32_t seed_base, int first, int count, ImbaRay* rays, void* states,
                    void* stream) {
    Frame f;
    int r = check_frame(frame, &f);
    if (r) return r;
    IMBA_CHECK_ARG(first >= 0 && count >= 0, "bad sizes");
    if (count == 0) return IMBA_OK;
    IMBA_CHECK_ARG(rays && states, "NULL buffer");
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    hipLaunchKernelGGL(k_vcm_camera, dim3(ceil_div(count, 256)), dim3(256), 0, s, f, seed_base, first, count,
                       (float4*)rays, (VcmState*)states);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_vcm_light_bounce(const ImbaVcmFrame* frame, const ImbaRay* rays, const ImbaHit* hits, void* states,
                          int capacity, const int* count_dev, void* vertices_a, void* vertices_b, ImbaRay* rays_out,
                          ImbaRay* shadow_rays, void* shadow_states, void* stream) {
    Frame f;
    int r = check_frame(frame, &f);
    if (r) return r;
    IMBA_CHECK_ARG(capacity >= 0, "capacity < 0");
    if (capacity == 0) return IMBA_OK;
    IMBA_CHECK_ARG(rays && hits && states && vertices_a && vertices_b && rays_out && shadow_rays && shadow_states,
                   "NULL buffer");
    LightArgs a;
    a.rays = (const float4*)rays;
    a.hits = (const int4*)hits;
    a.states = (VcmState*)states;
    a.rays_out = (float4*)rays_out;
    a.capacity = capacity;
    a.count_dev = count_dev;
    a.vtx_a = (float4*)vertices_a;
    a.vtx_b = (float4*)vertices_b;
    a.shadow_rays = (float4*)shadow_rays;
    a.shadow_states = (ShadowState*)shadow_states;
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    hipLaunchKernelGGL(k_vcm_light_bounce, dim3(ceil_div(capacity, 256)), dim3(256), 0, s, f, a);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_vcm_camera_bounce(const ImbaVcmFrame* frame, const ImbaRay* rays, const ImbaHit* hits, void* states,
                           int capacity, const int* count_dev, const void* vertices_a, const void* vertices_b,
                           const int* vertex_count_dev, ImbaRay* rays_out, ImbaRay* shadow_rays, void* shadow_states,
                           int* shadow_count_dev, void* queries, float* image, void* stream) {
    Frame f;
    int r = check_frame(frame, &f);
    if (r) return r;
    IMBA_CHECK_ARG(capacity >= 0, "capacity < 0");
    if (capacity == 0) return IMBA_OK;
    IMBA_CHECK_ARG(rays && hits && states && rays_out && shadow_rays && shadow_states && image, "NULL buffer");
    IMBA_CHECK_ARG(vertex_count_dev == nullptr || (vertices_a && vertices_b), "light vertices missing");
    CameraArgs a;
    a.rays = (const float4*)rays;
    a.hits = (const int4*)hits;
    a.states = (VcmState*)states;
    a.rays_out = (float4*)rays_out;
    a.capacity = capacity;
    a.count_dev = count_dev;
    a.image = image;
    a.vtx_a = (const float4*)vertices_a;
    a.vtx_b = (const float4*)vertices_b;
    a.vtx_count_dev = vertex_count_dev;
    a.shadow_rays = (float4*)shadow_rays;
    a.shadow_states = (ShadowState*)shadow_states;
    a.shadow_count_dev = shadow_count_dev;
    a.queries = (float4*)queries;
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    hipLaunchKernelGGL(k_vcm_camera_bounce, dim3(ceil_div(capacity, 256)), dim3(256), 0, s, f, a);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_vcm_merge_records(const ImbaHashGrid* grid, const void* vertices_a, const void* vertices_b, void* records,
                           void* stream) {
    IMBA_CHECK_ARG(grid && vertices_a && vertices_b && records, "NULL argument");
    const int n = grid->count;
    if (n == 0) return IMBA_OK;
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    hipLaunchKernelGGL(k_vcm_merge_records, dim3(ceil_div(n, 256)), dim3(256), 0, s, (const float4*)vertices_a,
                       (const float4*)vertices_b, grid->order, n, (MergeRec*)records);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_vcm_merge(const ImbaVcmFrame* frame, const ImbaHashGrid* grid, const void* records, const void* queries,
                   int capacity, const int* count_dev, int k, float* image, void* stream) {
    Frame f;
    int r = check_frame(frame, &f);
    if (r) return r;
    IMBA_CHECK_ARG(grid && records && queries && image, "NULL argument");
    IMBA_CHECK_ARG(k >= 1 && k <= 64, "the fused merge keeps its list in one wave: 1 <= k <= 64");
    IMBA_CHECK_ARG(capacity >= 0, "capacity < 0");
    if (capacity == 0 || grid->count == 0) return IMBA_OK;
    MergeArgs a;
    a.gp = grid->d_params;
    a.ends = grid->cell_ends;
    a.pos4 = grid->pos4;
    a.box = grid->bucket_box;
    a.rec = (const MergeRec*)records;
    a.queries = (const float4*)queries;
    a.capacity = capacity;
    a.count_dev = count_dev;
    a.k = k;
    a.albedo = frame->albedo;
    a.image = image;
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    hipLaunchKernelGGL(k_vcm_merge, dim3(ceil_div(capacity, 4)), dim3(256), 0, s, f, a);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

}  // extern "C"
