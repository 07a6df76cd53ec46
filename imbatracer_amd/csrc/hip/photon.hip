// Photon-mapping ray producers and vertex merging: the PPM mode of the VCM integrator
// (render/integrators/vcm.cpp with algo == ALGO_PPM), run on the device between traversals.
//
// Light paths (vcm.cpp:64-96, process_light_rays :196-256, bounce :129-193 with adjoint):
//   start  : path i is seeded as LightRayGen (render/ray_gen/ray_gen.h:138-145); a point on the
//            emitting triangle (sample_uniform_triangle, render/random.h:115-119) and a
//            cosine-weighted direction around its normal in the binormal/tangent/normal frame
//            (TriangleLight::sample_emit, render/light.h:109-141); throughput = intensity *
//            area * pi; tmin 1e-3.
//   at hit : a vertex on a non-specular material is stored as a photon (vcm.cpp:236-245) with
//            the throughput that reached it, the direction back along the ray and the path
//            length; then Russian roulette (render/random.h:183-188, from bounce rr_start on)
//            and a BSDF sample continue the path from the hit point with tmin = 1e-4 * t.
//            Paths end after max_depth hits (the reference ends light paths by roulette only).
//   photon slots: entry i of the bounce-b queue writes slot b * capacity + i, so the photons
//            compacted out of the slots are ordered by bounce, then by path (the reference
//            appends them from concurrent threads in no fixed order).
// Camera paths (process_camera_rays :372-441 with PPM: no light sampling, no connections, only
// specular lobes sampled): an emitter adds throughput * emission and ends the path; a
// non-specular hit records a merge query (position, normal, direction to the camera,
// throughput * albedo / pi) and ends the path; mirror and glass continue it.
// Merging (vertex_merging :572-611 with mis weight 1): for the k nearest photons within r,
// sum (1 - d^2 / r_k^2) * photon throughput over photons arriving on the query's side, where
// r_k^2 is the k-th distance when k photons were found and r^2 otherwise, times
// 2 / (pi r_k^2 light_paths) and the query's weight.
//
// Materials: kind 0 Lambertian (albedo), 1 mirror (reflectance = albedo), 2 glass (index of
// refraction 1.5, Fresnel-weighted choice between reflection and refraction, throughput
// unchanged).  Geometric normals are the triangles' own (outward for closed meshes).
#include <hip/hip_runtime.h>

#include "common.h"
#include "grid_query.h"
#include "sampling.h"

namespace imba {
namespace {

constexpr float kPi = 3.14159265359f;
constexpr float kFltMax = 3.402823466e38f;
constexpr float kGlassEta = 1.5f;

enum { kDiffuse = 0, kMirror = 1, kGlass = 2 };

struct LightArgs {
    float v0[3], v1[3], v2[3];
    float intensity[3];
    uint32_t seed_base;
    int first, count;
};

__global__ __launch_bounds__(256) void k_light(LightArgs a, float4* rays, PathState* states) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= a.count) return;
    const int i = a.first + j;
    Rng rng = seeded(a.seed_base, i, i);
    const Vec3 e1 = v3(a.v1[0] - a.v0[0], a.v1[1] - a.v0[1], a.v1[2] - a.v0[2]);
    const Vec3 e2 = v3(a.v2[0] - a.v0[0], a.v2[1] - a.v0[1], a.v2[2] - a.v0[2]);
    Vec3 n = v3(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
    const float area = 0.5f * sqrtf(dot3(n, n));
    n = normalize3(n);
    Vec3 tg, bn;
    frame(n, tg, bn);
    const float r1 = rng.uniform(), r2 = rng.uniform();
    const float su = sqrtf(r1);
    const float u = 1.0f - su, v = r2 * su, w = 1.0f - u - v;
    const Vec3 pos = v3(u * a.v0[0] + v * a.v1[0] + w * a.v2[0], u * a.v0[1] + v * a.v1[1] + w * a.v2[1],
                        u * a.v0[2] + v * a.v1[2] + w * a.v2[2]);
    const float u1 = rng.uniform(), u2 = rng.uniform();
    const Vec3 l = cos_hemisphere(u1, u2);
    const Vec3 d = v3(l.x * bn.x + l.y * tg.x + l.z * n.x, l.x * bn.y + l.y * tg.y + l.z * n.y,
                      l.x * bn.z + l.y * tg.z + l.z * n.z);
    rays[2 * (size_t)j] = make_float4(pos.x, pos.y, pos.z, 1e-3f);
    rays[2 * (size_t)j + 1] = make_float4(d.x, d.y, d.z, kFltMax);
    PathState s;
    s.pixel_id = i;
    s.sample_id = 0;
    s.rng = rng.s;
    const float scale = l.z / kPi > 0.0f ? area * kPi : 0.0f;  // zero pdf: no power
    s.tp[0] = a.intensity[0] * scale;
    s.tp[1] = a.intensity[1] * scale;
    s.tp[2] = a.intensity[2] * scale;
    s.bounces = 0;
    states[j] = s;
}

struct SurfArgs {
    const float4* rays;
    const int4* hits;
    PathState* states;
    float4* rays_out;      // may alias rays
    int count;             // entries (or, with count_dev, the capacity)
    const int* count_dev;  // device entry count read when the launch runs, or null
    const float* tri_normals;
    const int* tri_material;
    int n_tris;
    const float* albedo;    // 3 per material
    const float* emission;  // 3 per material (camera paths)
    const int* kind;        // per material
    int n_materials;
    int max_depth;
    int rr_start;
    // light paths
    float4* photon_rays;   // 2 float4 per slot: position, path length bits | direction to the previous vertex
    float4* photon_states; // 1 float4 per slot: path id bits (-1: empty), throughput
    // camera paths
    float4* queries;       // 4 float4 per sample
    int spp;
    float* image;
};

// Russian roulette (render/random.h:183-188); false ends the path.
__device__ __forceinline__ bool roulette(PathState& s, Rng& rng, int rr_start) {
    if (s.bounces < rr_start) return true;
    const float lum = 0.2126f * s.tp[0] + 0.7152f * s.tp[1] + 0.0722f * s.tp[2];
    const float p = fminf(1.0f, lum * 2.0f);
    if (!(rng.uniform() < p)) return false;
    const float ip = 1.0f / p;
    s.tp[0] *= ip; s.tp[1] *= ip; s.tp[2] *= ip;
    return true;
}

// Continues a path at a mirror or glass surface: returns the new direction.  nf faces the
// incoming ray's side; `entering` is true when the ray arrives on the normal's side.
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
    const float f = 0.5f * (rs * rs + rp * rp);
    if (rng.uniform() < f) return refl;
    return madd3(scale3(d, eta), nf, eta * cos_i - cos_t);
}

__device__ __forceinline__ void end_path(const SurfArgs& a, int i, PathState& s) {
    s.pixel_id = -1;
    a.states[i] = s;
}

__global__ __launch_bounds__(256) void k_light_bounce(SurfArgs a) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.count) return;
    const int n = a.count_dev ? min(max(*a.count_dev, 0), a.count) : a.count;
    a.photon_states[i] = make_float4(__int_as_float(-1), 0.0f, 0.0f, 0.0f);
    if (i >= n) return;
    PathState s = a.states[i];
    const int4 h = a.hits[i];
    if (h.x < 0 || h.x >= a.n_tris) return end_path(a, i, s);
    const int mat = a.tri_material[h.x];
    if (mat < 0 || mat >= a.n_materials) return end_path(a, i, s);
    const float4 o = a.rays[2 * (size_t)i];
    const float4 dr = a.rays[2 * (size_t)i + 1];
    const float t = __int_as_float(h.z);
    const Vec3 d = v3(dr.x, dr.y, dr.z);
    const Vec3 p = v3(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z);
    Vec3 nrm = v3(a.tri_normals[3 * (size_t)h.x], a.tri_normals[3 * (size_t)h.x + 1], a.tri_normals[3 * (size_t)h.x + 2]);
    const float cos_o = -dot3(d, nrm);
    if (cos_o == 0.0f) return end_path(a, i, s);
    const int kind = a.kind[mat];
    if (kind == kDiffuse) {
        a.photon_rays[2 * (size_t)i] = make_float4(p.x, p.y, p.z, __int_as_float(s.bounces + 2));
        a.photon_rays[2 * (size_t)i + 1] = make_float4(-d.x, -d.y, -d.z, 0.0f);
        a.photon_states[i] = make_float4(__int_as_float(s.pixel_id), s.tp[0], s.tp[1], s.tp[2]);
    }
    if (s.bounces + 1 >= a.max_depth) return end_path(a, i, s);
    Rng rng;
    rng.s = s.rng;
    if (!roulette(s, rng, a.rr_start)) return end_path(a, i, s);
    const bool entering = cos_o > 0.0f;
    const Vec3 nf = entering ? nrm : scale3(nrm, -1.0f);
    Vec3 w;
    if (kind == kDiffuse) {
        Vec3 tg, bn;
        frame(nf, tg, bn);
        const float u1 = rng.uniform(), u2 = rng.uniform();
        const Vec3 l = cos_hemisphere(u1, u2);
        w = v3(tg.x * l.x + bn.x * l.y + nf.x * l.z, tg.y * l.x + bn.y * l.y + nf.y * l.z,
               tg.z * l.x + bn.z * l.y + nf.z * l.z);
    } else {
        w = specular_dir(kind, d, nf, entering, rng);
    }
    if (kind != kGlass) {
        s.tp[0] *= a.albedo[3 * mat];
        s.tp[1] *= a.albedo[3 * mat + 1];
        s.tp[2] *= a.albedo[3 * mat + 2];
    }
    s.bounces++;
    s.rng = rng.s;
    a.states[i] = s;
    a.rays_out[2 * (size_t)i] = make_float4(p.x, p.y, p.z, 1e-4f * t);
    a.rays_out[2 * (size_t)i + 1] = make_float4(w.x, w.y, w.z, kFltMax);
}

__device__ __forceinline__ void end_camera_path(const SurfArgs& a, int i, PathState& s) {
    const size_t q = (size_t)s.pixel_id * a.spp + s.sample_id;
    a.queries[4 * q] = make_float4(__int_as_float(0x7f800000), __int_as_float(0x7f800000),
                                   __int_as_float(0x7f800000), __int_as_float(-1));
    end_path(a, i, s);
}

__global__ __launch_bounds__(256) void k_camera_bounce(SurfArgs a) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int n = a.count_dev ? min(max(*a.count_dev, 0), a.count) : a.count;
    if (i >= n) return;
    PathState s = a.states[i];
    const int4 h = a.hits[i];
    if (h.x < 0 || h.x >= a.n_tris) return end_camera_path(a, i, s);
    const int mat = a.tri_material[h.x];
    if (mat < 0 || mat >= a.n_materials) return end_camera_path(a, i, s);
    const float4 o = a.rays[2 * (size_t)i];
    const float4 dr = a.rays[2 * (size_t)i + 1];
    const float t = __int_as_float(h.z);
    const Vec3 d = v3(dr.x, dr.y, dr.z);
    const Vec3 p = v3(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z);
    Vec3 nrm = v3(a.tri_normals[3 * (size_t)h.x], a.tri_normals[3 * (size_t)h.x + 1], a.tri_normals[3 * (size_t)h.x + 2]);
    const float cos_o = -dot3(d, nrm);
    if (cos_o == 0.0f) return end_camera_path(a, i, s);
    const float er = a.emission[3 * mat], eg = a.emission[3 * mat + 1], eb = a.emission[3 * mat + 2];
    if (er > 0.0f || eg > 0.0f || eb > 0.0f) {
        if (a.image) {
            atomicAdd(&a.image[3 * (size_t)s.pixel_id + 0], s.tp[0] * er);
            atomicAdd(&a.image[3 * (size_t)s.pixel_id + 1], s.tp[1] * eg);
            atomicAdd(&a.image[3 * (size_t)s.pixel_id + 2], s.tp[2] * eb);
        }
        return end_camera_path(a, i, s);
    }
    if (s.bounces + 1 >= a.max_depth) return end_camera_path(a, i, s);
    const int kind = a.kind[mat];
    if (kind == kDiffuse) {
        const size_t q = (size_t)s.pixel_id * a.spp + s.sample_id;
        const float c = 1.0f / kPi;
        a.queries[4 * q] = make_float4(p.x, p.y, p.z, __int_as_float(s.pixel_id));
        a.queries[4 * q + 1] = make_float4(nrm.x, nrm.y, nrm.z, 0.0f);
        a.queries[4 * q + 2] = make_float4(-d.x, -d.y, -d.z, 0.0f);
        a.queries[4 * q + 3] = make_float4(s.tp[0] * a.albedo[3 * mat] * c, s.tp[1] * a.albedo[3 * mat + 1] * c,
                                           s.tp[2] * a.albedo[3 * mat + 2] * c, 0.0f);
        return end_path(a, i, s);
    }
    Rng rng;
    rng.s = s.rng;
    if (!roulette(s, rng, a.rr_start)) return end_camera_path(a, i, s);
    const bool entering = cos_o > 0.0f;
    const Vec3 nf = entering ? nrm : scale3(nrm, -1.0f);
    const Vec3 w = specular_dir(kind, d, nf, entering, rng);
    if (kind == kMirror) {
        s.tp[0] *= a.albedo[3 * mat];
        s.tp[1] *= a.albedo[3 * mat + 1];
        s.tp[2] *= a.albedo[3 * mat + 2];
    }
    s.bounces++;
    s.rng = rng.s;
    a.states[i] = s;
    a.rays_out[2 * (size_t)i] = make_float4(p.x, p.y, p.z, 1e-4f * t);
    a.rays_out[2 * (size_t)i + 1] = make_float4(w.x, w.y, w.z, kFltMax);
}

struct MergeArgs {
    const float4* queries;
    int nq, k;
    const int* out_idx;
    const float* out_dist;
    const int* out_count;
    const float4* photon_rays;
    const float4* photon_states;
    float radius_sqr;
    float inv_paths;
    float* image;
};

__global__ __launch_bounds__(256) void k_merge(MergeArgs a) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= a.nq) return;
    const int c = a.out_count[q];
    if (c <= 0) return;
    const float4 q0 = a.queries[4 * (size_t)q];
    const int pixel = __float_as_int(q0.w);
    if (pixel < 0) return;
    const float rs = c == a.k ? a.out_dist[(size_t)q * a.k + a.k - 1] : a.radius_sqr;
    if (!(rs > 0.0f)) return;
    const float4 qn = a.queries[4 * (size_t)q + 1];
    const float4 qo = a.queries[4 * (size_t)q + 2];
    const float4 qw = a.queries[4 * (size_t)q + 3];
    const float cos_c = qn.x * qo.x + qn.y * qo.y + qn.z * qo.z;
    float sr = 0.0f, sg = 0.0f, sb = 0.0f;
    for (int j = 0; j < c; j++) {
        const int pi = a.out_idx[(size_t)q * a.k + j];
        const float4 pd = a.photon_rays[2 * (size_t)pi + 1];
        const float cos_p = qn.x * pd.x + qn.y * pd.y + qn.z * pd.z;
        if (!(cos_p * cos_c > 0.0f)) continue;
        const float kern = 1.0f - a.out_dist[(size_t)q * a.k + j] / rs;
        const float4 ps = a.photon_states[pi];
        sr += kern * ps.y;
        sg += kern * ps.z;
        sb += kern * ps.w;
    }
    const float scale = 2.0f / (kPi * rs) * a.inv_paths;
    atomicAdd(&a.image[3 * (size_t)pixel + 0], qw.x * sr * scale);
    atomicAdd(&a.image[3 * (size_t)pixel + 1], qw.y * sg * scale);
    atomicAdd(&a.image[3 * (size_t)pixel + 2], qw.z * sb * scale);
}

// Merge records in grid slot order: the direction of photon order[s] and its state (weight
// in y, z, w), so that the fused merge reads them next to the positions the scan just read.
__global__ __launch_bounds__(256) void k_merge_records(const float4* photon_rays, const float4* photon_states,
                                                       const int* order, int n, float4* rec) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    const int v = order[s];
    rec[2 * (size_t)s] = photon_rays[2 * (size_t)v + 1];
    rec[2 * (size_t)s + 1] = photon_states[v];
}

struct GridMergeArgs {
    const GridParams* gp;
    const int* ends;
    const float4* pos4;
    const uint32_t* box;
    const float4* rec;
    const float4* queries;
    int nq, k;
    float radius_sqr;
    float inv_paths;
    float* image;
};

// The grid query and k_merge's sum in one launch, one query per wave: the wave's k-capped list
// keeps each photon's grid slot, lane j weighs list entry j, and the wave adds the lanes'
// terms in list order (k_merge's loop order).
__global__ __launch_bounds__(256) void k_merge_grid(GridMergeArgs a) {
    const int q = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (q >= a.nq) return;  // whole waves leave together
    const float4 q0 = a.queries[4 * (size_t)q];
    const int pixel = __float_as_int(q0.w);
    if (pixel < 0) return;
    float my_dist;
    int my_slot;
    const int c = query_wave_list<true>(*a.gp, a.ends, a.pos4, a.box, q0.x, q0.y, q0.z, a.k, my_dist, my_slot);
    if (c <= 0) return;
    const float kd = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(my_dist), a.k - 1));
    const float rs = c == a.k ? kd : a.radius_sqr;
    if (!(rs > 0.0f)) return;
    const float4 qn = a.queries[4 * (size_t)q + 1];
    const float4 qo = a.queries[4 * (size_t)q + 2];
    const float4 qw = a.queries[4 * (size_t)q + 3];
    const float cos_c = qn.x * qo.x + qn.y * qo.y + qn.z * qo.z;
    const int lane = (int)(threadIdx.x & 63);
    float cr = 0.0f, cg = 0.0f, cb = 0.0f;
    if (lane < c) {
        const float4 pd = a.rec[2 * (size_t)my_slot];
        const float cos_p = qn.x * pd.x + qn.y * pd.y + qn.z * pd.z;
        if (cos_p * cos_c > 0.0f) {
            const float kern = 1.0f - my_dist / rs;
            const float4 ps = a.rec[2 * (size_t)my_slot + 1];
            cr = kern * ps.y;
            cg = kern * ps.z;
            cb = kern * ps.w;
        }
    }
    float sr = 0.0f, sg = 0.0f, sb = 0.0f;
    for (int j = 0; j < c; j++) {
        sr += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cr), j));
        sg += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cg), j));
        sb += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cb), j));
    }
    if (lane == 0) {
        const float scale = 2.0f / (kPi * rs) * a.inv_paths;
        atomicAdd(&a.image[3 * (size_t)pixel + 0], qw.x * sr * scale);
        atomicAdd(&a.image[3 * (size_t)pixel + 1], qw.y * sg * scale);
        atomicAdd(&a.image[3 * (size_t)pixel + 2], qw.z * sb * scale);
    }
}

int launch_surface(bool camera, SurfArgs& a, void* stream) {
    if (a.count == 0) return IMBA_OK;
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    if (camera)
        hipLaunchKernelGGL(k_camera_bounce, dim3(ceil_div(a.count, 256)), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_light_bounce, dim3(ceil_div(a.count, 256)), dim3(256), 0, s, a);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

}  // namespace
}  // namespace imba

using namespace imba;

extern "C" {

int imba_ppm_light(const float* light_tri9, const float* intensity3, uint32_t seed_base, int first, int count,
                   ImbaRay* rays, void* states, void* stream) {
    IMBA_CHECK_ARG(light_tri9 && intensity3 && rays && states, "NULL argument");
    IMBA_CHECK_ARG(first >= 0 && count >= 0, "bad sizes");
    if (count == 0) return IMBA_OK;
    LightArgs a;
    for (int k = 0; k < 3; k++) {
        a.v0[k] = light_tri9[k];
        a.v1[k] = light_tri9[3 + k];
        a.v2[k] = light_tri9[6 + k];
        a.intensity[k] = intensity3[k];
    }
    a.seed_base = seed_base;
    a.first = first;
    a.count = count;
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    hipLaunchKernelGGL(k_light, dim3(ceil_div(count, 256)), dim3(256), 0, s, a, (float4*)rays, (PathState*)states);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_ppm_light_bounce(const ImbaRay* rays, const ImbaHit* hits, void* states, int count, const int* count_dev,
                          const float* tri_normals, const int* tri_materials, int tri_count, const float* albedo,
                          const int* kinds, int material_count, int max_depth, int rr_start, void* photon_rays,
                          void* photon_states, ImbaRay* rays_out, void* stream) {
    IMBA_CHECK_ARG(count >= 0, "count < 0");
    if (count == 0) return IMBA_OK;
    IMBA_CHECK_ARG(rays && hits && states && tri_normals && tri_materials && albedo && kinds && photon_rays &&
                       photon_states && rays_out,
                   "NULL argument");
    SurfArgs a = {};
    a.rays = (const float4*)rays;
    a.hits = (const int4*)hits;
    a.states = (PathState*)states;
    a.rays_out = (float4*)rays_out;
    a.count = count;
    a.count_dev = count_dev;
    a.tri_normals = tri_normals;
    a.tri_material = tri_materials;
    a.n_tris = tri_count;
    a.albedo = albedo;
    a.kind = kinds;
    a.n_materials = material_count;
    a.max_depth = max_depth;
    a.rr_start = rr_start;
    a.photon_rays = (float4*)photon_rays;
    a.photon_states = (float4*)photon_states;
    return launch_surface(false, a, stream);
}

int imba_ppm_camera_bounce(const ImbaRay* rays, const ImbaHit* hits, void* states, int count, const int* count_dev,
                           const float* tri_normals, const int* tri_materials, int tri_count, const float* albedo,
                           const float* emission, const int* kinds, int material_count, int max_depth, int rr_start,
                           int spp, void* queries, float* image, ImbaRay* rays_out, void* stream) {
    IMBA_CHECK_ARG(count >= 0 && spp > 0, "bad sizes");
    if (count == 0) return IMBA_OK;
    IMBA_CHECK_ARG(rays && hits && states && tri_normals && tri_materials && albedo && emission && kinds && queries &&
                       rays_out,
                   "NULL argument");
    SurfArgs a = {};
    a.rays = (const float4*)rays;
    a.hits = (const int4*)hits;
    a.states = (PathState*)states;
    a.rays_out = (float4*)rays_out;
    a.count = count;
    a.count_dev = count_dev;
    a.tri_normals = tri_normals;
    a.tri_material = tri_materials;
    a.n_tris = tri_count;
    a.albedo = albedo;
    a.emission = emission;
    a.kind = kinds;
    a.n_materials = material_count;
    a.max_depth = max_depth;
    a.rr_start = rr_start;
    a.queries = (float4*)queries;
    a.spp = spp;
    a.image = image;
    return launch_surface(true, a, stream);
}

int imba_ppm_merge(const void* queries, int nq, int k, const int* out_idx, const float* out_dist, const int* out_count,
                   const void* photon_rays, const void* photon_states, float radius, int light_paths, float* image,
                   void* stream) {
    IMBA_CHECK_ARG(nq >= 0 && k >= 1 && light_paths > 0, "bad sizes");
    if (nq == 0) return IMBA_OK;
    IMBA_CHECK_ARG(queries && out_idx && out_dist && out_count && photon_rays && photon_states && image,
                   "NULL argument");
    MergeArgs a;
    a.queries = (const float4*)queries;
    a.nq = nq;
    a.k = k;
    a.out_idx = out_idx;
    a.out_dist = out_dist;
    a.out_count = out_count;
    a.photon_rays = (const float4*)photon_rays;
    a.photon_states = (const float4*)photon_states;
    a.radius_sqr = radius * radius;
    a.inv_paths = 1.0f / (float)light_paths;
    a.image = image;
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    hipLaunchKernelGGL(k_merge, dim3(ceil_div(nq, 256)), dim3(256), 0, s, a);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_ppm_merge_records(const ImbaHashGrid* grid, const void* photon_rays, const void* photon_states,
                           void* records, void* stream) {
    IMBA_CHECK_ARG(grid && photon_rays && photon_states && records, "NULL argument");
    const int n = grid->count;
    if (n == 0) return IMBA_OK;
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    hipLaunchKernelGGL(k_merge_records, dim3(ceil_div(n, 256)), dim3(256), 0, s, (const float4*)photon_rays,
                       (const float4*)photon_states, grid->order, n, (float4*)records);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_ppm_merge_grid(const ImbaHashGrid* grid, const void* records, const void* queries, int nq, int k,
                        float radius, int light_paths, float* image, void* stream) {
    IMBA_CHECK_ARG(grid, "grid is NULL");
    IMBA_CHECK_ARG(nq >= 0 && light_paths > 0, "bad sizes");
    IMBA_CHECK_ARG(k >= 1 && k <= 64, "the fused merge keeps its list in one wave: 1 <= k <= 64");
    if (nq == 0 || grid->count == 0) return IMBA_OK;
    IMBA_CHECK_ARG(records && queries && image, "NULL argument");
    GridMergeArgs a;
    a.gp = grid->d_params;
    a.ends = grid->cell_ends;
    a.pos4 = grid->pos4;
    a.box = grid->bucket_box;
    a.rec = (const float4*)records;
    a.queries = (const float4*)queries;
    a.nq = nq;
    a.k = k;
    a.radius_sqr = radius * radius;
    a.inv_paths = 1.0f / (float)light_paths;
    a.image = image;
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    hipLaunchKernelGGL(k_merge_grid, dim3(ceil_div(nq, 4)), dim3(256), 0, s, a);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

}  // extern "C"
