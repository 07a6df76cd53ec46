// Path-tracing ray producers for the measured workloads: camera rays and Lambertian bounces
// written straight into device ray queues, so a frame's bounce loop never leaves the GPU.
//
// Camera rays follow PathTracer's sample function (render/integrators/pt.cpp:168-177) with the
// per-sample seeding of PixelRayGen::fill_queue (render/ray_gen/ray_gen.h:51-72) and the
// pinhole camera of render/ray_gen/camera.h:57-79.  A bounce follows PathTracer::bounce
// (pt.cpp:46-80) for a Lambertian material: terminate past the maximum path length, Russian
// roulette (render/random.h:183-188) from bounce `rr_start` on, cosine-weighted direction
// (random.h:80-83) in the frame of random.h:57-72 around the geometric normal, throughput
// times the albedo.  Rays that hit an emissive material add throughput * emission to the
// image and end.  Terminated rays get pixel_id = -1 (integrator.h:125-128) and are removed by
// imba_compact_rays.
#include <hip/hip_runtime.h>

#include "common.h"
#include "pt_bounce.h"
#include "sampling.h"

namespace imba {
namespace {

struct CamArgs {
    float m[16];  // raster -> world (row major), applied to (raster y, raster x, 0, 1)
    float pos[3];
    int width, height, spp;
    uint32_t seed_base;
    int first;  // first sample index of this launch
    int count;
    float4* queries;  // photon-mapping merge queries (4 float4 per sample) to mark empty, or null
};

__global__ __launch_bounds__(256) void k_camera(CamArgs c, float4* rays, PtState* states) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= c.count) return;
    const int i = c.first + j;
    const int pixel = i / c.spp;
    const int sample = i % c.spp;
    const int y = pixel / c.width, x = pixel % c.width;
    Rng rng = seeded(c.seed_base, i, pixel);
    const float sx = (float)x + rng.uniform();
    const float sy = (float)y + rng.uniform();
    float t[4];
    for (int r = 0; r < 4; r++) t[r] = c.m[4 * r] * sy + c.m[4 * r + 1] * sx + c.m[4 * r + 2] * 0.0f + c.m[4 * r + 3];
    float dx = t[0] / t[3] - c.pos[0], dy = t[1] / t[3] - c.pos[1], dz = t[2] / t[3] - c.pos[2];
    const float inv = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
    dx *= inv; dy *= inv; dz *= inv;
    rays[2 * (size_t)j] = make_float4(c.pos[0], c.pos[1], c.pos[2], 0.0f);
    rays[2 * (size_t)j + 1] = make_float4(dx, dy, dz, 3.402823466e38f);
    PtState s;
    s.pixel_id = pixel;
    s.sample_id = sample;
    s.rng = rng.s;
    s.tp[0] = s.tp[1] = s.tp[2] = 1.0f;
    s.bounces = 0;
    states[j] = s;
    if (c.queries)
        c.queries[4 * (size_t)i] = make_float4(__int_as_float(0x7f800000), __int_as_float(0x7f800000),
                                               __int_as_float(0x7f800000), __int_as_float(-1));
}

__global__ __launch_bounds__(256) void k_bounce(BounceArgs a) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int n = a.count_dev ? min(max(*a.count_dev, 0), a.count) : a.count;
    if (i >= n) return;
    bounce_entry(a, i);
}

}  // namespace
}  // namespace imba

using namespace imba;

extern "C" {

static int camera(const float* raster_to_world16, const float* cam_pos3, int width, int height, int spp,
                  uint32_t seed_base, int first, int count, ImbaRay* rays, void* states, void* queries, void* stream) {
    IMBA_CHECK_ARG(raster_to_world16 && cam_pos3 && rays && states, "NULL argument");
    IMBA_CHECK_ARG(width > 0 && height > 0 && spp > 0 && first >= 0 && count >= 0, "bad sizes");
    IMBA_CHECK_ARG((int64_t)first + count <= (int64_t)width * height * spp, "samples out of range");
    if (count == 0) return IMBA_OK;
    CamArgs c;
    for (int k = 0; k < 16; k++) c.m[k] = raster_to_world16[k];
    for (int k = 0; k < 3; k++) c.pos[k] = cam_pos3[k];
    c.width = width;
    c.height = height;
    c.spp = spp;
    c.seed_base = seed_base;
    c.first = first;
    c.count = count;
    c.queries = (float4*)queries;
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    hipLaunchKernelGGL(k_camera, dim3(ceil_div(count, 256)), dim3(256), 0, s, c, (float4*)rays, (PtState*)states);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_pt_camera(const float* raster_to_world16, const float* cam_pos3, int width, int height, int spp,
                   uint32_t seed_base, int first, int count, ImbaRay* rays, void* states, void* stream) {
    return camera(raster_to_world16, cam_pos3, width, height, spp, seed_base, first, count, rays, states, nullptr,
                  stream);
}

int imba_ppm_camera(const float* raster_to_world16, const float* cam_pos3, int width, int height, int spp,
                    uint32_t seed_base, int first, int count, ImbaRay* rays, void* states, void* queries,
                    void* stream) {
    IMBA_CHECK_ARG(queries, "NULL argument");
    return camera(raster_to_world16, cam_pos3, width, height, spp, seed_base, first, count, rays, states, queries,
                  stream);
}

static int pt_bounce(const ImbaRay* rays, const ImbaHit* hits, void* states, int count, const int* count_dev,
                     const float* tri_normals, const int* tri_materials, int tri_count, const float* albedo,
                     const float* emission, int material_count, int max_path_len, int rr_start, float offset,
                     float* image, ImbaRay* rays_out, void* stream) {
    IMBA_CHECK_ARG(count >= 0, "count < 0");
    if (count == 0) return IMBA_OK;
    IMBA_CHECK_ARG(rays && hits && states && tri_normals && tri_materials && albedo && emission && rays_out,
                   "NULL argument");
    BounceArgs a;
    a.rays = (const float4*)rays;
    a.hits = (const int4*)hits;
    a.states = (PtState*)states;
    a.rays_out = (float4*)rays_out;
    a.count = count;
    a.count_dev = count_dev;
    a.tri_normals = tri_normals;
    a.tri_material = tri_materials;
    a.n_tris = tri_count;
    a.albedo = albedo;
    a.emission = emission;
    a.n_materials = material_count;
    a.max_path_len = max_path_len;
    a.rr_start = rr_start;
    a.offset = offset;
    a.image = image;
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    hipLaunchKernelGGL(k_bounce, dim3(ceil_div(count, 256)), dim3(256), 0, s, a);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_pt_bounce(const ImbaRay* rays, const ImbaHit* hits, void* states, int count, const float* tri_normals,
                   const int* tri_materials, int tri_count, const float* albedo, const float* emission,
                   int material_count, int max_path_len, int rr_start, float offset, float* image, ImbaRay* rays_out,
                   void* stream) {
    return pt_bounce(rays, hits, states, count, nullptr, tri_normals, tri_materials, tri_count, albedo, emission,
                     material_count, max_path_len, rr_start, offset, image, rays_out, stream);
}

int imba_pt_bounce_dev(const ImbaRay* rays, const ImbaHit* hits, void* states, int capacity, const int* count_dev,
                       const float* tri_normals, const int* tri_materials, int tri_count, const float* albedo,
                       const float* emission, int material_count, int max_path_len, int rr_start, float offset,
                       float* image, ImbaRay* rays_out, void* stream) {
    IMBA_CHECK_ARG(count_dev, "count_dev is NULL");
    return pt_bounce(rays, hits, states, capacity, count_dev, tri_normals, tri_materials, tri_count, albedo, emission,
                     material_count, max_path_len, rr_start, offset, image, rays_out, stream);
}

}  // extern "C"
