/*
 * traversal_gpu.h — drop-in for the AnyDSL-generated GPU traversal header that the reference
 * includes as `namespace traversal_gpu { #include <traversal_gpu.h> }`
 * (src/imbatracer/core/traversal_interface.h:8-10).
 *
 * Type and field names are the ones the reference's host code uses:
 *   Node.left_bb / right_bb .lo_x .. .hi_z, left, right     core/gpu_adapter.cpp:62-103, 14-31
 *   Vec4 x y z w, Vec2 x y                                  core/gpu_adapter.cpp:132-141, scene.cpp:84-86
 *   InstanceNode transf / id / next / pad[2]                core/gpu_adapter.cpp:285-295
 *   Ray org / dir (Vec4), Hit tri_id / inst_id / tmax / u   render/integrators/integrator.h:72-90
 *   TransparencyMask = MaskBuffer::MaskDesc (3 ints)        core/mask.h:9-20, scene.cpp:122-124
 * The byte layout is that of include/imba_hip.h (same sizes and offsets).
 *
 * The two functions are implemented by libimba_hip.so (HIP, gfx950).  As in the reference
 * (render/scheduling/ray_queue.h:297-308, 347-358) every pointer is device memory, `count` rays
 * are traversed, and the call returns when the hits are written.  Errors are reported on
 * stderr (the reference's functions return nothing).  Array sizes are those of the allocations
 * holding the pointers at the time of the call, and nothing derived from the arrays is kept
 * between calls, so the caller may rewrite or reallocate them between calls (imba_trace_desc
 * in imba_hip.h describes the traversal).
 */
#ifndef IMBA_TRAVERSAL_GPU_H
#define IMBA_TRAVERSAL_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct Vec4 { float x, y, z, w; } Vec4;
typedef struct Vec2 { float x, y; } Vec2;
typedef struct BBox { float lo_x, hi_x, lo_y, hi_y, lo_z, hi_z; } BBox;
typedef struct Node { BBox left_bb; BBox right_bb; int left; int right; int pad[2]; } Node;
typedef struct InstanceNode { Vec4 transf[3]; int id; int next; int pad[2]; } InstanceNode;
typedef struct Ray { Vec4 org; Vec4 dir; } Ray;
typedef struct Hit { int tri_id; int inst_id; float tmax; float u; } Hit;
typedef struct TransparencyMask { int width; int height; int offset; } TransparencyMask;

/* Closest hit of rays[0..count) (ray_queue.h:297-308). */
void intersect_gpu_masked_instanced(int root, Node* nodes, InstanceNode* instances, Vec4* tris, Ray* rays,
                                    Hit* hits, int* indices, Vec2* texcoords, TransparencyMask* masks,
                                    char* mask_buffer, int count);
/* Any hit (shadow rays) of rays[0..count) (ray_queue.h:347-358). */
void occluded_gpu_masked_instanced(int root, Node* nodes, InstanceNode* instances, Vec4* tris, Ray* rays,
                                   Hit* hits, int* indices, Vec2* texcoords, TransparencyMask* masks,
                                   char* mask_buffer, int count);

#ifdef __cplusplus
}
#endif

#endif /* IMBA_TRAVERSAL_GPU_H */
