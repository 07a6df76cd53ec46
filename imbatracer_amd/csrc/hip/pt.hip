// Path-tracing ray producers for the measured workloads: camera rays, the shading step of every
// hit (next-event shadow ray, emitter contribution, Lambertian bounce) and the shadow-ray
// contributions, written straight into device ray queues, so a frame's bounce loop never
// leaves the GPU.
//
// Camera rays follow PathTracer's sample function (render/integrators/pt.cpp:168-177) with the
// per-sample seeding of PixelRayGen::fill_queue (render/ray_gen/ray_gen.h:51-72) and the
// pinhole camera of render/ray_gen/camera.h:57-79.  The shading step is pt_shade_entry
// (pt_shade.h): PathTracer::process_primary_rays for Lambertian and emitting triangles.
// Terminated paths get pixel_id = -1 (integrator.h:125-128) and are removed by
// imba_compact_rays; empty shadow slots have pixel_id = -1 in their shadow state.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "common.h"
#include "pt_shade.h"
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

// S: PtState (path tracer, 40 bytes) or PathState (photon mapping, 32 bytes).
template <typename S>
__global__ __launch_bounds__(256) void k_camera(CamArgs c, float4* rays, S* states) {
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
    S s;
    s.pixel_id = pixel;
    s.sample_id = sample;
    s.rng = rng.s;
    s.tp[0] = s.tp[1] = s.tp[2] = 1.0f;
    s.bounces = 0;
    if constexpr (std::is_same<S, PtState>::value) {
        s.last_pdf = 0.0f;  // read only after a bounce (bounces > 0)
        s.last_specular = 0;
    }
    states[j] = s;
    if (c.queries)
        c.queries[4 * (size_t)i] = make_float4(__int_as_float(0x7f800000), __int_as_float(0x7f800000),
                                               __int_as_float(0x7f800000), __int_as_float(-1));
}

__device__ __forceinline__ int entries(int capacity, const int* count_dev) {
    return count_dev ? min(max(*count_dev, 0), capacity) : capacity;
}

__global__ __launch_bounds__(256) void k_pt_shade(PtShadeArgs a, int capacity, const int* count_dev) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= entries(capacity, count_dev)) return;
    pt_shade_entry(a, i);
}

__global__ __launch_bounds__(256) void k_pt_shadow(const int4* hits, const ShadowState* st, float* image, int capacity,
                                                   const int* count_dev) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= entries(capacity, count_dev)) return;
    pt_shadow_entry(hits, st, image, i);
}

}  // namespace

// Checks the shading description (host side of the device pointers it holds).
int check_shading(const ImbaPtShading* sh) {
    IMBA_CHECK_ARG(sh, "shading is NULL");
    IMBA_CHECK_ARG(sh->tri_normals && sh->tri_materials && sh->albedo && sh->emission && sh->emitter_area,
                   "shading: NULL array");
    IMBA_CHECK_ARG(sh->tri_count >= 0 && sh->material_count >= 0, "shading: negative count");
    IMBA_CHECK_ARG(sh->light_count >= 0 && (sh->light_count == 0 || sh->lights), "shading: lights");
    IMBA_CHECK_ARG(sh->max_path_len >= 1 && sh->rr_start >= 0, "shading: max_path_len >= 1 and rr_start >= 0");
    return IMBA_OK;
}

}  // namespace imba

using namespace imba;

extern "C" {

static int camera(const float* raster_to_world16, const float* cam_pos3, int width, int height, int spp,
                  uint32_t seed_base, int first, int count, ImbaRay* rays, void* states, void* queries, bool pt,
                  void* stream) {
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
    if (pt)
        hipLaunchKernelGGL(k_camera<PtState>, dim3(ceil_div(count, 256)), dim3(256), 0, s, c, (float4*)rays,
                           (PtState*)states);
    else
        hipLaunchKernelGGL(k_camera<PathState>, dim3(ceil_div(count, 256)), dim3(256), 0, s, c, (float4*)rays,
                           (PathState*)states);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_pt_camera(const float* raster_to_world16, const float* cam_pos3, int width, int height, int spp,
                   uint32_t seed_base, int first, int count, ImbaRay* rays, void* states, void* stream) {
    return camera(raster_to_world16, cam_pos3, width, height, spp, seed_base, first, count, rays, states, nullptr,
                  true, stream);
}

int imba_ppm_camera(const float* raster_to_world16, const float* cam_pos3, int width, int height, int spp,
                    uint32_t seed_base, int first, int count, ImbaRay* rays, void* states, void* queries,
                    void* stream) {
    IMBA_CHECK_ARG(queries, "NULL argument");
    return camera(raster_to_world16, cam_pos3, width, height, spp, seed_base, first, count, rays, states, queries,
                  false, stream);
}

int imba_pt_shade(const ImbaPtShading* sh, const ImbaRay* rays, const ImbaHit* hits, void* states, int capacity,
                  const int* count_dev, float* image, ImbaRay* rays_out, ImbaRay* shadow_rays, void* shadow_states,
                  void* stream) {
    int r = check_shading(sh);
    if (r) return r;
    IMBA_CHECK_ARG(capacity >= 0, "capacity < 0");
    if (capacity == 0) return IMBA_OK;
    IMBA_CHECK_ARG(rays && hits && states && rays_out, "NULL argument");
    IMBA_CHECK_ARG((shadow_rays == nullptr) == (shadow_states == nullptr), "shadow_rays and shadow_states go together");
    PtShadeArgs a;
    a.rays = (const float4*)rays;
    a.hits = (const int4*)hits;
    a.states = (PtState*)states;
    a.rays_out = (float4*)rays_out;
    a.shadow_rays = (float4*)shadow_rays;
    a.shadow_states = (ShadowState*)shadow_states;
    a.image = image;
    a.sh = *sh;
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    hipLaunchKernelGGL(k_pt_shade, dim3(ceil_div(capacity, 256)), dim3(256), 0, s, a, capacity, count_dev);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_pt_shadow(const ImbaHit* shadow_hits, const void* shadow_states, int capacity, const int* count_dev,
                   float* image, void* stream) {
    IMBA_CHECK_ARG(capacity >= 0, "capacity < 0");
    if (capacity == 0) return IMBA_OK;
    IMBA_CHECK_ARG(shadow_hits && shadow_states && image, "NULL argument");
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    hipLaunchKernelGGL(k_pt_shadow, dim3(ceil_div(capacity, 256)), dim3(256), 0, s, (const int4*)shadow_hits,
                       (const ShadowState*)shadow_states, image, capacity, count_dev);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

}  // extern "C"
