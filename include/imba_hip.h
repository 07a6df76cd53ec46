/*
 * imba_hip.h — C ABI of the MI355X (gfx950) traversal / compaction / hash-grid library.
 *
 * Every entry point is extern "C", takes plain pointers and sizes and returns an int status
 * (IMBA_OK == 0).  On failure imba_last_error() returns a message for the calling thread.
 * Pointers documented as "device" must be device memory of the current HIP device
 * (imba_malloc or any hipMalloc'd memory); "host" pointers are ordinary host memory.
 * A `stream` argument is a hipStream_t passed as void*; NULL means the library's stream for
 * the calling thread, and such calls return only after the work has finished.
 *
 * The reference's own FFI for traversal (the AnyDSL-generated functions
 * {intersect,occluded}_gpu_masked_instanced) is declared with the reference's type names in
 * traversal_gpu.h, which is implemented by the same library.
 */
#ifndef IMBA_HIP_H
#define IMBA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------------------------
 * Traversal ABI types.  Field names follow the reference's traversal ABI (re-exported at
 * src/imbatracer/core/traversal_interface.h:12-17, fields used at
 * render/scheduling/ray_queue.h:42-54, core/gpu_adapter.cpp:14-31,62-141,276-295,
 * render/integrators/integrator.h:72-90, core/mask.h:9-20).  Byte layout is this library's
 * own and is fixed here (sizes are checked with static_asserts in the library).
 * ---------------------------------------------------------------------------------------- */
typedef struct ImbaVec4 { float x, y, z, w; } ImbaVec4;                 /* 16 B */
typedef struct ImbaVec2 { float x, y; } ImbaVec2;                       /*  8 B */

/* org.w = tmin, dir.w = tmax (camera.h:61-64, pt.cpp:38-41). 32 B. */
typedef struct ImbaRay { ImbaVec4 org; ImbaVec4 dir; } ImbaRay;

/* tri_id < 0: no hit / unoccluded.  tri_id is the global triangle id written in the
 * triangle records, inst_id the instance index, tmax the hit distance in the ray's own
 * parameterisation, u the barycentric weight of the triangle's second vertex. 16 B. */
typedef struct ImbaHit { int tri_id; int inst_id; float tmax; float u; } ImbaHit;

typedef struct ImbaBBox { float lo_x, hi_x, lo_y, hi_y, lo_z, hi_z; } ImbaBBox;   /* 24 B */

/* BVH2 node.  left/right >= 0: node index; < 0: ~(index of the first record of a leaf)
 * (first Vec4 of the leaf's triangles in a mesh BVH, first InstanceNode in the top level).
 * right == IMBA_DUMMY_CHILD marks the dummy parent written above a single-leaf tree. 64 B,
 * so that one node is four 16-byte loads. */
typedef struct ImbaNode {
    ImbaBBox left_bb;
    ImbaBBox right_bb;
    int left;
    int right;
    int pad[2];
} ImbaNode;

/* Instance record: transf is the world->object 3x4 matrix (rows), id the instance index,
 * next the root node of the instanced mesh BVH; pad = {-1,-1} ends a top-level leaf. 64 B. */
typedef struct ImbaInstanceNode {
    ImbaVec4 transf[3];
    int id;
    int next;
    int pad[2];
} ImbaInstanceNode;

/* Opacity mask of one material: width x height bytes starting at mask_buffer[offset]. */
typedef struct ImbaTransparencyMask { int width; int height; int offset; } ImbaTransparencyMask;

#define IMBA_DUMMY_CHILD 0x76543210
#define IMBA_LEAF_END_BITS 0x80000000u   /* bit pattern of v2.w on the last triangle of a leaf */

/* ------------------------------------------------------------------------------------------ */
enum {
    IMBA_OK = 0,
    IMBA_ERR_INVALID_ARG = 1,
    IMBA_ERR_HIP = 2,
    IMBA_ERR_OUT_OF_MEMORY = 3,
    IMBA_ERR_BAD_BVH = 4,
    IMBA_ERR_CAPACITY = 5,
    IMBA_ERR_NO_DEVICE = 6,
    IMBA_ERR_KERNEL = 7
};

/* Message of the last failed call on this thread ("" if none). */
const char* imba_last_error(void);
/* Library version string. */
const char* imba_version(void);

/* ---------------------------- device memory, streams, events -----------------------------
 * Replacement for the anydsl::Array / anydsl::copy runtime used by the reference
 * (ray_queue.h:77-80,277,292; scene.cpp:14-27,121-151). */
int imba_device_count(int* count);
int imba_set_device(int device);
int imba_get_device(int* device);   /* the calling thread's current device */
int imba_device_info(int* cu_count, int* lds_bytes_per_block, int* l2_bytes, char* name, int name_len);
int imba_malloc(void** ptr, size_t bytes);
int imba_free(void* ptr);
int imba_host_alloc(void** ptr, size_t bytes);     /* pinned host memory */
int imba_host_free(void* ptr);
int imba_memcpy_h2d(void* dst_device, const void* src_host, size_t bytes, void* stream);
int imba_memcpy_d2h(void* dst_host, const void* src_device, size_t bytes, void* stream);
int imba_memcpy_d2d(void* dst_device, const void* src_device, size_t bytes, void* stream);
/* Byte fill of device memory (hipMemsetAsync: not for code captured into a graph). */
int imba_memset(void* dst_device, int value, size_t bytes, void* stream);
int imba_stream_create(void** stream);
int imba_stream_destroy(void* stream);
int imba_stream_sync(void* stream);
int imba_event_create(void** event);
int imba_event_destroy(void* event);
int imba_event_record(void* event, void* stream);
int imba_event_elapsed_ms(float* ms, void* start, void* end);
int imba_event_sync(void* event);  /* waits until the work before the event's record is done */
/* Work enqueued on stream after this call waits for the work recorded before event (a device
 * side wait: the host does not block). */
int imba_stream_wait_event(void* stream, void* event);

/* ------------------------------ host BVH construction ------------------------------------
 * Restatement of the reference builders and GPU layout writers:
 *   SplitBvhBuilder (core/sbvh_builder.h:28-165), FastBvhBuilder (core/fast_bvh_builder.h:43-160),
 *   GpuMeshAdapter / GpuTopLevelAdapter (core/gpu_adapter.cpp:33-308),
 *   Scene::build_mesh_accels / build_top_level_accel (render/scene.cpp:37-117). */
enum { IMBA_BUILDER_SBVH = 0, IMBA_BUILDER_BINNED_SAH = 1 };

typedef struct ImbaMeshDesc {
    const float* vertices;      /* host, 4 floats per vertex (x, y, z, w) */
    int vertex_count;
    const uint32_t* indices;    /* host, 4 per triangle: i0, i1, i2, material */
    int tri_count;
    const float* texcoords;     /* host, 2 floats per vertex, or NULL (all zero) */
} ImbaMeshDesc;

/* Host-side traversal arrays (the TraversalData of ray_queue.h:42-54 before upload). */
typedef struct ImbaSceneDesc {
    int root;
    const ImbaNode* nodes;                 int node_count;
    const ImbaInstanceNode* instances;     int instance_count;
    const ImbaVec4* tris;                  int tri_vec4_count;
    const int* indices;                    int index_count;      /* 4 per triangle */
    const ImbaVec2* texcoords;             int texcoord_count;
    const ImbaTransparencyMask* masks;     int mask_count;
    const uint8_t* mask_buffer;            int mask_buffer_size;
} ImbaSceneDesc;

typedef struct ImbaSceneBuild ImbaSceneBuild;

int imba_build_create(ImbaSceneBuild** out);
int imba_build_release(ImbaSceneBuild* b);
/* Copies the mesh. mesh_id is its index (order of calls). */
int imba_build_add_mesh(ImbaSceneBuild* b, const ImbaMeshDesc* mesh, int* mesh_id);
/* mat: row-major 4x4 object->world; the inverse is computed as core/float4x4.h:144-191. */
int imba_build_add_instance(ImbaSceneBuild* b, int mesh_id, const float* mat16, int* inst_id);
/* Optional opacity mask for a material (bytes != 0 are opaque). Materials without one
 * are opaque (MaskBuffer default, core/mask.h:22-26,43-45). */
int imba_build_set_mask(ImbaSceneBuild* b, int material, int width, int height, const uint8_t* texels);
/* Builds every mesh BVH with `builder`, then the top level with the binned-SAH builder.
 * threads > 1 builds meshes in parallel; the output is identical for any thread count. */
int imba_build_run(ImbaSceneBuild* b, int builder, int threads);
/* Host pointers stay valid until imba_build_release. */
int imba_build_get_desc(const ImbaSceneBuild* b, ImbaSceneDesc* out);
/* Per-global-triangle geometric normal (3 floats) and material, for shading kernels. */
int imba_build_get_tri_attrs(const ImbaSceneBuild* b, const float** geom_normals, const int** materials, int* tri_count);

/* CPU traversal layout, the arrays the reference's CPU traversal reads
 * (intersect_cpu_masked_instanced, ray_queue.h:254-265), written as CpuMeshAdapter and
 * CpuTopLevelAdapter write them (core/cpu_adapter.cpp:14-206, 221-366):
 *   - a node has four child slots: box j and children[j].  children[j] > 0 is a node index
 *     (children always follow their parent, so 0 never names a child); < 0 is ~(first Vec4 of
 *     the leaf's packets) at mesh level and ~(first instance record of the leaf) at top level;
 *     0 is an unused slot, whose box is [FLT_MAX, -FLT_MAX] (all zero in the dummy parent of a
 *     tree that is a single leaf).  Top-level nodes follow the mesh nodes; root = their first.
 *   - a leaf is a run of 13-Vec4 packets of up to four triangles, one per lane:
 *     v0.x, v0.y, v0.z, e1.x, e1.y, e1.z, e2.x, e2.y, e2.z, n.x, n.y, n.z, id, with
 *     e1 = v0 - v1, e2 = v2 - v0, n = e1 x e2 (fma form: n.x = fma(e1.y, e2.z, -(e1.z * e2.y))
 *     and cyclic) and id the global triangle id as int bits; an unused lane has zero
 *     geometry and id INT_MIN.  After a leaf's packets comes one Vec4 whose four lanes are
 *     0x80000000.
 *   - instance records as in the GPU layout, in the order of the top-level BVH4's leaves. */
typedef struct ImbaNode4 {
    float min_x[4], max_x[4];
    float min_y[4], max_y[4];
    float min_z[4], max_z[4];
    int children[4];
    int pad[4];
} ImbaNode4;                                  /* 128 B: two cache lines */

typedef struct ImbaCpuSceneDesc {
    int root;
    const ImbaNode4* nodes;                int node_count;
    const ImbaInstanceNode* instances;     int instance_count;
    const ImbaVec4* tris;                  int tri_vec4_count;
    const int* indices;                    int index_count;
    const ImbaVec2* texcoords;             int texcoord_count;
    const ImbaTransparencyMask* masks;     int mask_count;
    const uint8_t* mask_buffer;            int mask_buffer_size;
} ImbaCpuSceneDesc;

/* Builds the CPU layout of the scene last built by imba_build_run: mesh BVH4s with `builder`
 * (SBVH with leaf threshold 2 and alpha 1e-4 as CpuMeshAdapter, cpu_adapter.cpp:37; or binned
 * SAH), the top-level BVH4 with the binned-SAH builder (cpu_adapter.cpp:243-247), costs
 * ((count - 1) / 4 + 1) * area and 0.5 * area (cpu_adapter.cpp:45-52). */
int imba_build_cpu_layout(ImbaSceneBuild* b, int builder, int threads);
/* Host pointers stay valid until imba_build_release or the next imba_build_cpu_layout. */
int imba_build_get_cpu_desc(const ImbaSceneBuild* b, ImbaCpuSceneDesc* out);

/* Acceleration-structure cache files (the container of loaders/load_bvh.cpp): u32 magic
 * 0x313F1A57, then blocks {u64 size, u32 type, u32 node_count, u32 prim_count, node records,
 * Vec4 records}; type IMBA_CACHE_BVH2 (1) holds a mesh's BVH2 layout (ImbaNode, 3 Vec4 per
 * triangle), IMBA_CACHE_BVH4 (2) its CPU layout (ImbaNode4, triangle packets).  A block stores
 * one tree with links relative to its first node / first Vec4 and triangle ids relative to the
 * mesh's first id.  store rebases by -offsets and refuses a file that already holds a block of
 * the type; load rebases by +offsets (load_bvh.cpp:66-95, 121-139).  With nodes_out or tris_out
 * NULL or too small, load returns IMBA_ERR_CAPACITY with the counts filled in. */
enum { IMBA_CACHE_BVH2 = 1, IMBA_CACHE_BVH4 = 2 };
int imba_bvh_store(const char* path, int type, const void* nodes, int node_count, const ImbaVec4* tris,
                   int vec4_count, int node_offset, int tris_offset, int tri_id_offset);
int imba_bvh_load(const char* path, int type, int node_offset, int tris_offset, int tri_id_offset,
                  void* nodes_out, int node_cap, ImbaVec4* tris_out, int vec4_cap, int* node_count,
                  int* vec4_count);
/* 1 if the file holds a block of `type`, else 0. */
int imba_bvh_has_block(const char* path, int type);
/* Per-mesh cache file for imba_build_run / imba_build_cpu_layout (Scene::build_mesh_accels,
 * render/scene.cpp:49-70): a mesh whose file holds the layout's block is loaded instead of
 * built; a built mesh is stored to its file.  path NULL or "" removes the file. */
int imba_build_set_cache(ImbaSceneBuild* b, int mesh_id, const char* path);
/* Meshes loaded from cache files / built by the last imba_build_run or imba_build_cpu_layout. */
int imba_build_cache_stats(const ImbaSceneBuild* b, int* loaded, int* built);

/* Runs one builder and writes the sequence of nodes and leaves it emits (for checking the
 * builders and for custom layout writers).  Records, as int32 (floats as their bits):
 *   node: 0, count, parent box, count child boxes      leaf: 1, count, box, count primitive ids
 * A box is min.x, min.y, min.z, max.x, max.y, max.z.  width is 2 or 4 children per node;
 * cost selects the cost functions of the reference adapters.  *len receives the number of
 * int32 written, or needed when cap is too small (then IMBA_ERR_CAPACITY is returned). */
enum { IMBA_COST_GPU_MESH = 0, IMBA_COST_GPU_TOP = 1, IMBA_COST_CPU = 2 };
int imba_bvh_emit_mesh(const ImbaMeshDesc* mesh, int builder, int width, int cost,
                       int32_t* out, size_t cap, size_t* len);
/* Binned SAH over given primitive boxes (6 floats) and centres (3 floats). */
int imba_bvh_emit_boxes(const float* boxes, const float* centres, int n, int width, int cost,
                        int32_t* out, size_t cap, size_t* len);

/* The 4-wide tree the upload derives from a host scene (for inspection): every node holds up to
 * four children of a BVH2 node and its absorbed inner children, chosen largest surface area
 * first; leaves and their boxes are the BVH2's.  Nodes: ImbaNode4, children[j] > 0 node, < 0 the
 * BVH2's leaf link, 0 unused; instance records with `next` = 4-wide mesh root (-1 for records
 * the top level does not reach).  Pointers stay valid until imba_wide_release. */
typedef struct ImbaWideBvh ImbaWideBvh;
int imba_wide_create(const ImbaSceneDesc* host, ImbaWideBvh** out);
int imba_wide_get(const ImbaWideBvh* w, const ImbaNode4** nodes, int* node_count,
                  const ImbaInstanceNode** instances, int* instance_count, int* root, int* max_stack);
/* The same tree with quantized 64-byte nodes (node_count 0 if a box is not finite): origin[3]
 * (float), exponent[3] (int8) and child count (uint8), children[4] (int), then uint8
 * lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4] and 8 bytes of padding; child box j on
 * an axis is [fma(lo[j], 2^e, origin), fma(hi[j], 2^e, origin)], which contains the exact box. */
int imba_wide_get_quantized(const ImbaWideBvh* w, const void** nodes, int* node_count);
int imba_wide_release(ImbaWideBvh* w);

/* ------------------------------ device-resident scenes -----------------------------------
 * Scene::upload_* (render/scene.cpp:119-188).  The scene is validated on upload: child
 * indices in range and increasing from parent to child, leaf records in range, instance
 * roots in range, leaf sentinels present. */
typedef struct ImbaScene ImbaScene;
/* The upload's checks alone, on host arrays (no device needed): IMBA_ERR_BAD_BVH with a
 * message for a malformed scene; otherwise the deepest traversal stack a ray can need and
 * whether every mask texel is opaque. */
int imba_scene_validate(const ImbaSceneDesc* host, int* max_stack, int* all_opaque);
int imba_scene_upload(const ImbaSceneDesc* host, ImbaScene** out);
int imba_scene_release(ImbaScene* s);
/* Device pointers of the resident arrays (counts as uploaded). */
int imba_scene_device_desc(const ImbaScene* s, ImbaSceneDesc* out);
/* Bytes of device memory held by the scene. */
int imba_scene_footprint(const ImbaScene* s, size_t* bytes);

/* ------------------------------ traversal -------------------------------------------------
 * Closest hit (K1) and any hit (K2) of `count` rays (device) against a resident scene,
 * hits written to device memory.  Semantics of one ray:
 *   - a triangle is hit at t when the ray meets it inside or on its edges, det != 0,
 *     tmin <= t and t <= tmax, and its material mask is opaque at the hit's texcoords;
 *   - closest hit: smallest t; equal t -> smallest tri_id, then smallest inst_id;
 *   - occluded: the first hit found in traversal order (near child first, ties left);
 *   - no hit: tri_id = inst_id = -1, tmax = ray tmax, u = 0. */
typedef struct ImbaTraceConfig {
    int block_size;        /* threads per block: 64..1024, multiple of 64 (0 = default)    */
    int stack_lds;         /* per-lane stack entries kept in LDS (0 = default)             */
    int persistent;        /* 1: persistent waves fetching rays with a global counter      */
    int blocks_per_cu;     /* grid of this many blocks per CU, threads striding over the
                              rays (0 = default: persistent launches fill the resident
                              blocks, others launch one thread per ray)                 */
    int bvh_width;         /* 2: the scene's BVH2 as given; 4: the 4-wide tree derived from
                              it at upload (same leaves); 0 = default                      */
    int quantized;         /* with bvh_width 4: 64-byte nodes whose child boxes are 8-bit
                              offsets from the node box, rounded outward                  */
    int treelet;           /* quantized tree only: nodes of its top levels (0..256) each
                              block copies into LDS and reads from there (0 = none);
                              batches keep one copy per wave, of at most 64 nodes        */
} ImbaTraceConfig;

int imba_intersect(const ImbaScene* s, const ImbaRay* rays, ImbaHit* hits, int count,
                   const ImbaTraceConfig* cfg, void* stream);
int imba_occluded(const ImbaScene* s, const ImbaRay* rays, ImbaHit* hits, int count,
                  const ImbaTraceConfig* cfg, void* stream);
/* Device-count variants (capturable in a graph): the queue holds *count_dev rays (clamped to
 * [0, capacity]) when the launch runs; the grid is sized for `capacity`. */
int imba_intersect_dev(const ImbaScene* s, const ImbaRay* rays, ImbaHit* hits, int capacity, const int* count_dev,
                       const ImbaTraceConfig* cfg, void* stream);
int imba_occluded_dev(const ImbaScene* s, const ImbaRay* rays, ImbaHit* hits, int capacity, const int* count_dev,
                      const ImbaTraceConfig* cfg, void* stream);
/* Same traversal, also writes per-ray visit counts (device, 3 uint32 per ray:
 * inner nodes whose children were box-tested, triangles tested, instances entered). */
int imba_intersect_stats(const ImbaScene* s, const ImbaRay* rays, ImbaHit* hits, uint32_t* counts,
                         int count, void* stream);
int imba_occluded_stats(const ImbaScene* s, const ImbaRay* rays, ImbaHit* hits, uint32_t* counts,
                        int count, void* stream);
/* Visit counts of the tree a configuration traverses (cfg->bvh_width; the *_stats calls above
 * count on the BVH2, as the CPU restatement does): nodes whose children were box-tested
 * (4-wide nodes for width 4), triangles tested, instances entered. */
int imba_trace_stats(const ImbaScene* s, const ImbaRay* rays, ImbaHit* hits, uint32_t* counts, int count,
                     int any_hit, const ImbaTraceConfig* cfg, void* stream);
/* Stack use of the tree a configuration traverses: adds one to depth_hist[d] (device, 64 uint32
 * bins, not cleared here) for every ray whose deepest traversal stack entry was d (bin 63: 63
 * and deeper).  With the 16-entry LDS stack of the default launches, the rays of bins > 16
 * spill stack entries to global memory.  Same hits as imba_intersect / imba_occluded. */
int imba_trace_stack_depths(const ImbaScene* s, const ImbaRay* rays, ImbaHit* hits, int count, int any_hit,
                            const ImbaTraceConfig* cfg, uint32_t* depth_hist, void* stream);
/* Traversal over caller-owned device arrays described by `dev` (device pointers and
 * element counts; a count of 0 means "as many records as the allocation holding the pointer
 * has room for"), for callers that manage their own TraversalData.  Nothing derived from the
 * arrays is kept between calls, so the caller may rewrite or reallocate them at any time:
 * every call traverses the BVH2 as given, with the mask test on (unless indices, masks or
 * mask_buffer are empty), a stack of 64 entries per ray and every index bounds-checked.
 * any_hit: 0 closest hit, 1 occluded.  This is also the path of the reference FFI entry points
 * (traversal_gpu.h). */
int imba_trace_desc(const ImbaSceneDesc* dev, const ImbaRay* rays, ImbaHit* hits, int count, int any_hit,
                    void* stream);
/* Rays of imba_trace_desc / FFI calls on the current device that met a malformed structure or
 * needed more than 64 stack entries (cumulative). */
int imba_ffi_error_count(int* count);

/* Batched traversal of many queues (the loop of RayQueue::traverse_gpu / traverse_occluded_gpu
 * calls that QueueScheduler::process_queues and the TileScheduler threads issue,
 * render/scheduling/queue_scheduler.h:193-218, tile_scheduler.h:120-132), one launch per
 * kernel variant (closest / occluded x opaque / masked scene).  Every queue's hits equal those
 * of its own imba_intersect / imba_occluded call.  The batch keeps the queue table, work
 * counters and spill stacks in device memory of its own: a launch allocates nothing (after a
 * first launch of a shape that needs spill stacks) and may be captured into a graph
 * (imba_graph_*).  Launches of one batch must be ordered (one stream). */
typedef struct ImbaTraceItem {
    const ImbaScene* scene;
    const ImbaRay* rays;       /* device */
    ImbaHit* hits;             /* device */
    int count;                 /* rays; with count_dev, the most rays *count_dev may hold */
    const int* count_dev;      /* NULL, or a device int with the ray count, read when the launch runs */
    int any_hit;               /* 0: closest hit, 1: occluded */
} ImbaTraceItem;
typedef struct ImbaTraceBatch ImbaTraceBatch;
int imba_trace_batch_create(const ImbaTraceItem* items, int n_items, ImbaTraceBatch** out);
/* Replaces the batch's queues (same rules as create).  The new table is copied on `stream`
 * (pinned staging, no device-wide synchronisation unless the batch must grow): launch on the
 * same stream after it.  For request loops whose set of queues changes every step. */
int imba_trace_batch_update(ImbaTraceBatch* b, const ImbaTraceItem* items, int n_items, void* stream);
/* cfg: block_size / stack_lds / blocks_per_cu as for imba_intersect; every launch is persistent. */
int imba_trace_batch_launch(ImbaTraceBatch* b, const ImbaTraceConfig* cfg, void* stream);
int imba_trace_batch_release(ImbaTraceBatch* b);

/* Launch shape autotuning: times every supported shape (static and persistent) of every tree
 * (BVH2, 4-wide, quantized with LDS treelets of 0 / 64 / 256 nodes) on the given device queue
 * with HIP events, `reps` launches each after one warm-up, and returns the fastest in *best
 * (and its mean time per launch in *best_ms, nullable).  Results do not depend on the shape. */
int imba_autotune(const ImbaScene* s, const ImbaRay* rays, ImbaHit* hits, int count, int any_hit, int reps,
                  ImbaTraceConfig* best, float* best_ms, void* stream);

/* ------------------------------ graphs ----------------------------------------------------
 * Capture of a sequence of this library's asynchronous calls on one stream into a replayable
 * HIP graph (the bounce loop of a frame, for instance).  Calls to be captured take a created
 * stream (not NULL), device-side counts (the *_dev variants) and workspaces sized
 * beforehand; a call that would synchronize or allocate fails during capture. */
typedef struct ImbaGraph ImbaGraph;
int imba_graph_begin(void* stream);
int imba_graph_end(void* stream, ImbaGraph** out);
int imba_graph_launch(ImbaGraph* g, void* stream);
int imba_graph_release(ImbaGraph* g);
/* Number of rays the last launches found a malformed BVH for (0 for a valid scene). */
int imba_trace_error_count(const ImbaScene* s, int* count);

/* ------------------------------ ray-queue compaction --------------------------------------
 * RayQueue::compact_hits (ray_queue.h:161-185): moves rays, hits and states of rays with
 * hits[i].tri_id >= 0 to the front, keeping their order; the misses follow, also in order.
 * RayQueue::compact_rays (ray_queue.h:188-207): moves rays and states whose state starts
 * with an int pixel_id >= 0 to the front, keeping their order; hits are not moved.
 * Out-of-place: inputs and outputs are distinct device buffers.  *new_count (host) receives
 * the number of hits / survivors; for compact_rays it is `count` when no ray was removed.
 * state_stride is the size in bytes of one state record (multiple of 4). */
int imba_compact_hits(const ImbaRay* rays_in, const ImbaHit* hits_in, const void* states_in,
                      int state_stride, int count,
                      ImbaRay* rays_out, ImbaHit* hits_out, void* states_out,
                      int* new_count, void* stream);
int imba_compact_rays(const ImbaRay* rays_in, const void* states_in, int state_stride, int count,
                      ImbaRay* rays_out, void* states_out, int* new_count, void* stream);
/* Allocation-free compact_hits.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, hosts that forbid
 * hipMalloc).  Same arguments, results and errors as imba_compact_hits, except:
 *  - the block offsets live in a library array of 256 KiB + 4 B per device, held until the
 *    process ends, so count may not exceed IMBA_COMPACT_NOALLOC_MAX_COUNT (134,217,728;
 *    larger counts return IMBA_ERR_CAPACITY);
 *  - every call on one device shares that array: calls that overlap in time on one device
 *    (other host threads, other streams) race and give wrong results;
 *  - stream NULL is the device's null stream (no per-thread stream is created);
 *  - new_count may be NULL: nothing is read back and, with a stream given, nothing waits.
 * It calls no hipMalloc / hipFree. */
#define IMBA_COMPACT_NOALLOC_MAX_COUNT 134217728
int imba_compact_hits_noalloc(const ImbaRay* rays_in, const ImbaHit* hits_in, const void* states_in,
                              int state_stride, int count,
                              ImbaRay* rays_out, ImbaHit* hits_out, void* states_out,
                              int* new_count, void* stream);
/* Device-count variant of compact_rays for graph-captured loops: the queue holds *count_dev
 * (clamped to [0, capacity]) entries when the launch runs and the new size is written to the
 * device int *new_count_dev; nothing is read back to the host.  `workspace` is device memory
 * of at least imba_compact_workspace_size(capacity) bytes. */
size_t imba_compact_workspace_size(int capacity);
int imba_compact_rays_dev(const ImbaRay* rays_in, const void* states_in, int state_stride, int capacity,
                          const int* count_dev, ImbaRay* rays_out, void* states_out, int* new_count_dev,
                          void* workspace, void* stream);

/* RayQueue::sort_by_material (ray_queue.h:211-244) on the device: for the first `count`
 * entries (the hits, after compact_hits), the material of the hit triangle, indices[4 tri_id + 3]
 * (the lookup of pt.cpp:85-91), is written to rays[i].dir.w and `sorted` receives the entry
 * indices grouped by material in increasing material order, and in queue order within a
 * material (the reference's parallel scatter leaves that order open).  Entries whose triangle
 * or material is out of range (a miss, for instance) get key num_mats and come last.
 * bucket_start (num_mats + 1 ints, nullable): first position of each material and of that last
 * group.  indices: device, 4 ints per triangle, tri_count triangles. */
int imba_sort_by_material(const int* indices, int tri_count, ImbaRay* rays, const ImbaHit* hits, int count,
                          int num_mats, int* sorted, int* bucket_start, void* stream);

/* ------------------------------ photon hash grid ------------------------------------------
 * HashGrid::build / query (rangesearch/rangesearch.h:25-148).  Photons are records of
 * `stride` bytes whose position (3 floats) is at byte `pos_offset`.  The grid keeps the
 * photons' indices ordered by cell, and within a cell by index.  The bucket table only
 * grows (max(previous size, 2 * count)), as in the reference. */
typedef struct ImbaHashGrid ImbaHashGrid;
typedef struct ImbaHashGridInfo {
    int photon_count;
    int table_size;                 /* number of buckets                                   */
    float bbox_min[3], bbox_max[3]; /* photon bounds expanded by 0.1% of the extents       */
    float radius, radius_sqr, cell_size, inv_cell_size;
    const int* cell_ends;           /* device, table_size entries: end of each bucket      */
    const int* order;               /* device, photon_count entries: photon index per slot */
    const float* positions;         /* device, 4 floats per slot: x, y, z, index bits      */
} ImbaHashGridInfo;

int imba_hashgrid_create(ImbaHashGrid** out);
int imba_hashgrid_release(ImbaHashGrid* g);
int imba_hashgrid_build(ImbaHashGrid* g, const void* photons, int stride, int pos_offset,
                        int count, float radius, void* stream);
/* For each query position (records of query_stride bytes, position at query_offset): up to
 * k photon indices (out_idx[q*k + j]) sorted by squared distance (out_dist, may be NULL),
 * with the reference's rules for duplicates and ties; out_count[q] = number written.
 * 1 <= k <= 4096 (k <= 64: the list lives in the lanes of one wave per query; above, in LDS,
 * k * 8 bytes per query; the reference sizes its list by k, rangesearch.h:98). */
int imba_hashgrid_query(const ImbaHashGrid* g, const void* queries, int query_stride, int query_offset,
                        int nq, int k, int* out_idx, float* out_dist, int* out_count, void* stream);
int imba_hashgrid_info(const ImbaHashGrid* g, ImbaHashGridInfo* info);
/* Forgets the table-size history (the next build sizes its table as a first build would) and
 * keeps the allocations: one grid object serves independent photon sets, e.g. one request
 * after another, without reallocating.  No counterpart in the reference, whose grid lives as
 * long as its integrator. */
int imba_hashgrid_reset(ImbaHashGrid* g);
/* Copies photon records into grid order (the reference's photons_ array): dst[slot] = src[order[slot]]. */
int imba_hashgrid_gather(const ImbaHashGrid* g, const void* src, int stride, void* dst, void* stream);

/* ------------------------------ path-tracing ray producers --------------------------------
 * Device-side producers of the queues a path tracer traverses: the shading that the reference
 * runs on the host around RayQueue (render/integrators/pt.cpp), for the materials of the
 * measured scenes (Lambertian surfaces, triangle emitters) and triangle lights.
 * States are 40-byte records (PTState): int pixel_id, int sample_id, uint64 rng,
 * float throughput[3], int bounces, float last_pdf, int last_specular.  Shadow states are
 * 16-byte records: int pixel_id (-1: empty slot), float contribution[3].
 * imba_pt_camera: samples first..first+count-1 of a width x height x spp image; sample i is
 *   pixel i / spp, seeded as render/ray_gen/ray_gen.h:51-72 with seed_base, jittered as
 *   pt.cpp:168-177 and shot through raster_to_world (imba_gen_camera_matrix).
 * imba_pt_shade: PathTracer::process_primary_rays (pt.cpp:82-160) for every entry of the queue
 *   and its hit: a miss ends the path; an emitter adds throughput * radiance * MIS weight to
 *   image and ends the path; otherwise next-event estimation (pt.cpp:17-43) writes the shadow
 *   ray of entry i to slot i of shadow_rays / shadow_states (both NULL: none), and a Lambertian
 *   bounce (pt.cpp:46-80) writes the continuation ray to rays_out (may equal rays) with
 *   tmin = 1e-3 t.  Russian roulette draws its number on every bounce and applies from bounce
 *   rr_start on (0: every bounce, as the reference).  Ended paths get pixel_id = -1.
 * imba_pt_shadow: process_shadow_rays (integrator.h:51-66): every non-empty slot whose shadow
 *   hit has tri_id < 0 adds its contribution to image.
 * count_dev (nullable): the queue size read on the device when the launch runs (clamped to
 * [0, capacity]); without it `capacity` entries. */
typedef struct ImbaPtShading {
    const float* tri_normals;   /* device, 3 per global triangle: the geometric normal
                                   normalize(cross(v1 - v0, v2 - v0)), also the shading normal */
    const int* tri_materials;   /* device, material of each global triangle                  */
    int tri_count;
    const float* albedo;        /* device, 3 per material: Lambertian reflectance            */
    const float* emission;      /* device, 3 per material: emitted radiance (0: no emitter)  */
    const float* emitter_area;  /* device, 1 per material: area of its emitting triangle     */
    int material_count;
    const float* lights;        /* device, 12 per triangle light: p0, p1, p2, intensity
                                   (TriangleLight, render/light.h:102-170)                  */
    int light_count;            /* 0: no next-event estimation                               */
    int max_path_len;
    int rr_start;
} ImbaPtShading;

int imba_pt_camera(const float* raster_to_world16, const float* cam_pos3, int width, int height, int spp,
                   uint32_t seed_base, int first, int count, ImbaRay* rays, void* states, void* stream);
int imba_pt_shade(const ImbaPtShading* shading, const ImbaRay* rays, const ImbaHit* hits, void* states,
                  int capacity, const int* count_dev, float* image, ImbaRay* rays_out, ImbaRay* shadow_rays,
                  void* shadow_states, void* stream);
int imba_pt_shadow(const ImbaHit* shadow_hits, const void* shadow_states, int capacity, const int* count_dev,
                   float* image, void* stream);

/* ------------------------------ batched shading + compaction ------------------------------
 * One path-tracing shading step (imba_pt_shade) and the stable compaction of the survivors
 * (imba_compact_rays, state stride 40) for many queues at once, in five launches whatever the
 * number of queues: the queues' device counts are read when the launch runs, so a bounce loop
 * over many queues and scenes can be captured in one graph (nothing in it allocates or sets
 * memory).  For every item the result equals imba_pt_shade + imba_compact_rays_dev on that
 * queue alone: rays/states are shaded in place, survivors go to rays_out/states_out in queue
 * order and their number to *next_count_dev (when non-null), shadow rays to the item's shadow
 * slots.  imba_shade_batch_shadow then adds every item's unoccluded shadow contributions
 * (imba_pt_shadow over shadow_hits, same entry counts as the preceding launch).  The
 * reference's schedulers shade and compact each queue on the host (queue_scheduler.h:177-275,
 * ray_queue.h:188-207). */
typedef struct ImbaShadeItem {
    ImbaRay* rays;             /* the queue (shaded in place)                                 */
    const ImbaHit* hits;
    void* states;              /* 40-byte path-tracing states                                */
    int capacity;              /* entries the buffers hold                                   */
    const int* count_dev;      /* device count of the queue (NULL: capacity)                 */
    ImbaRay* rays_out;         /* compacted survivors                                        */
    void* states_out;
    int* next_count_dev;       /* device count of the survivors (nullable)                   */
    ImbaRay* shadow_rays;      /* shadow slots, one per entry (NULL: no next-event rays)     */
    void* shadow_states;
    const ImbaHit* shadow_hits;/* the shadow slots' any-hit results (imba_shade_batch_shadow) */
    float* image;              /* nullable                                                   */
    ImbaPtShading shading;     /* the queue's scene                                          */
} ImbaShadeItem;

typedef struct ImbaShadeBatch ImbaShadeBatch;
int imba_shade_batch_create(const ImbaShadeItem* items, int count, ImbaShadeBatch** out);
/* Replaces the items (table copied through pinned memory on `stream`; the host array may be
 * reused when the call returns). */
int imba_shade_batch_update(ImbaShadeBatch* b, const ImbaShadeItem* items, int count, void* stream);
int imba_shade_batch_launch(ImbaShadeBatch* b, void* stream);
int imba_shade_batch_shadow(ImbaShadeBatch* b, void* stream);
int imba_shade_batch_release(ImbaShadeBatch* b);

/* ------------------------------ photon-mapping ray producers ------------------------------
 * The PPM mode of the VCM integrator (render/integrators/vcm.cpp, algo == ALGO_PPM) on the
 * device: light paths deposit photons that HashGrid indexes, camera paths end in merge queries
 * answered by imba_hashgrid_query.  States are 32-byte records: int pixel_id (the light path
 * index for light paths), int sample_id, uint64 rng, float throughput[3], int bounces.  Material kinds: 0 Lambertian, 1 mirror,
 * 2 glass (index of refraction 1.5).
 * imba_ppm_light: light paths first..first+count-1 from the triangle light_tri9 (3 vertices;
 *   it emits around cross(v1 - v0, v2 - v0)), seeded as LightRayGen
 *   (render/ray_gen/ray_gen.h:122-152), sampled as TriangleLight::sample_emit
 *   (render/light.h:109-141); throughput = intensity * area * pi.
 * imba_ppm_light_bounce: one hit of every light path of the queue (process_light_rays,
 *   vcm.cpp:196-256): a non-specular hit stores a photon in slot i of photon_rays (2 float4:
 *   position, path length | direction back along the ray) and photon_states (1 float4: path
 *   index, throughput); every other slot of the count entries (or of the capacity, with
 *   count_dev) gets path index -1.  Paths end after max_depth hits, by Russian roulette from
 *   bounce rr_start on, or on a miss; the next ray starts at the hit with tmin = 1e-4 t.
 *   Photon slots are stable-compacted with imba_compact_rays (state stride 16).
 * imba_ppm_camera: imba_pt_camera, and marks the merge query of each sample empty.
 * imba_ppm_camera_bounce: one hit of every camera path (process_camera_rays, vcm.cpp:372-441,
 *   PPM): an emitter adds throughput * emission to image and ends the path; a non-specular hit
 *   writes the sample's merge query (4 float4: position, pixel | normal | direction to the
 *   camera | throughput * albedo / pi) and ends the path; mirror and glass continue it.
 * imba_ppm_merge: vertex_merging (vcm.cpp:572-611, mis weight 1) from imba_hashgrid_query's
 *   output over the queries: Epanechnikov-weighted photon throughput into image.
 * imba_ppm_merge_records + imba_ppm_merge_grid: the same merge without the index lists: the
 *   photons' directions and states gathered once into grid slot order (32 bytes per photon,
 *   the grid built over photon_rays), then the grid query and the merge in one launch that
 *   reads them by slot (1 <= k <= 64).  The lists, distances, terms and per-query sums are
 *   those of imba_hashgrid_query + imba_ppm_merge (same order of operations).
 * count_dev (nullable): the queue size read on the device when the launch runs. */
int imba_ppm_light(const float* light_tri9, const float* intensity3, uint32_t seed_base, int first, int count,
                   ImbaRay* rays, void* states, void* stream);
int imba_ppm_light_bounce(const ImbaRay* rays, const ImbaHit* hits, void* states, int count, const int* count_dev,
                          const float* tri_normals, const int* tri_materials, int tri_count, const float* albedo,
                          const int* kinds, int material_count, int max_depth, int rr_start, void* photon_rays,
                          void* photon_states, ImbaRay* rays_out, void* stream);
int imba_ppm_camera(const float* raster_to_world16, const float* cam_pos3, int width, int height, int spp,
                    uint32_t seed_base, int first, int count, ImbaRay* rays, void* states, void* queries,
                    void* stream);
int imba_ppm_camera_bounce(const ImbaRay* rays, const ImbaHit* hits, void* states, int count, const int* count_dev,
                           const float* tri_normals, const int* tri_materials, int tri_count, const float* albedo,
                           const float* emission, const int* kinds, int material_count, int max_depth, int rr_start,
                           int spp, void* queries, float* image, ImbaRay* rays_out, void* stream);
int imba_ppm_merge(const void* queries, int nq, int k, const int* out_idx, const float* out_dist, const int* out_count,
                   const void* photon_rays, const void* photon_states, float radius, int light_paths, float* image,
                   void* stream);
int imba_ppm_merge_records(const ImbaHashGrid* grid, const void* photon_rays, const void* photon_states,
                           void* records, void* stream);
int imba_ppm_merge_grid(const ImbaHashGrid* grid, const void* records, const void* queries, int nq, int k,
                        float radius, int light_paths, float* image, void* stream);

/* ------------------------------ vertex connection and merging ---------------------------
 * One iteration of VCMIntegrator<algo> (render/integrators/vcm.cpp) on the device, for the
 * materials of the measured scenes (kinds: 0 Lambertian, 1 mirror, 2 glass of index 1.5) and
 * triangle lights, with the reference's partial MIS weights (dVC, dVCM, dVM) and float
 * operation order.  States are IMBA_VCM_STATE_BYTES-byte records: int pixel_id (light paths:
 * the path index; -1 once ended), int sample_id, uint64 rng, float throughput[3],
 * int path_length, float dVC, dVCM, dVM, int pad.  Shadow states are the 16-byte records of
 * imba_pt_shadow (int pixel_id, -1 for an empty slot; float contribution[3]), whose
 * unoccluded contributions imba_pt_shadow adds to the image.
 * Light vertices (LightPathVertex, light_vertices.h:17-37) are two 32-byte records each:
 *   vertices_a: float4(position, path_length bits), float4(direction back along the light
 *               ray, triangle id bits);
 *   vertices_b: float4(key bits, throughput) with key = path index or -1 for an empty slot,
 *               float4(dVCM, dVM, dVC, material bits).
 * imba_vcm_light: light paths first..first+count-1 (trace_light_paths, vcm.cpp:64-96): path i
 *   belongs to light (runs of light_paths / light_count, the remainder to light 0), seeded as
 *   LightRayGen (ray_gen.h:138-145), sampled as TriangleLight::sample_emit (light.h:109-141).
 * imba_vcm_light_bounce: process_light_rays (vcm.cpp:196-256) for every entry: the partial
 *   weights at the hit, a light vertex into slot i of vertices_a / vertices_b (every other slot
 *   of the capacity gets key -1) and the camera connection (connect_to_camera :258-318) into
 *   shadow slot i on a Lambertian hit, then the bounce (:129-193); paths end after
 *   light_max_depth hits (the reference ends light paths by Russian roulette only).
 *   The vertex slots of all bounces are stable-compacted with imba_compact_rays (vertices_a as
 *   the rays, vertices_b as 32-byte states) into the dense light vertex cache.
 * imba_vcm_camera: camera samples first..first+count-1 (trace_camera_paths :98-127), seeded as
 *   PixelRayGen, dVCM = light_paths / pdf_cam_w.
 * imba_vcm_camera_bounce: process_camera_rays (vcm.cpp:320-443) for every entry: emitter hits
 *   with their MIS weight into image, direct_illum (:445-488) into shadow slot (1 + C) i,
 *   num_connections = C connections to light vertices drawn from the cache (:490-570) into
 *   slots (1 + C) i + 1 .. (1 + C) i + C, the merge query of a Lambertian vertex into query
 *   slot i (4 float4: position | pixel bits, normal | dVCM, direction to the previous vertex |
 *   dVM, throughput | material bits; an empty slot has position +inf and pixel -1), then the
 *   bounce.  *shadow_count_dev (nullable) receives (1 + C) * entries.  vertex_count_dev: the
 *   device count of the cache (NULL: no connections).
 * imba_vcm_merge_records: 32-byte merge records of the cache in the grid's slot order (the
 *   grid built over vertices_a positions, stride 32): direction | dVCM, throughput | dVM.
 * imba_vcm_merge: vertex_merging (vcm.cpp:572-611) for every query slot: the k nearest light
 *   vertices within the radius (the grid's query, 1 <= k <= 64) and their Epanechnikov- and
 *   MIS-weighted throughputs into image, in one launch that reads the records by grid slot.
 * count_dev (nullable): the queue size read on the device when the launch runs.
 * algo (VCMSubAlgorithm, vcm.h:37-43) switches techniques as the reference's template does:
 *   VCM  every technique below;
 *   BPT  no merging: eta_vm = 0 in every partial weight (vcm.cpp:45), no merge queries (:427);
 *        the caller builds the cache but no grid (:98-99);
 *   LT   light paths only (:47-51): no vertex is stored (:238), the camera connections carry
 *        MIS weight 1 (:301); the caller runs no camera pass;
 *   PT   camera paths only: next event with MIS weight 1 (:473), no connections (:424), no
 *        merging; emitter hits keep their MIS weight (:404); the caller runs no light pass.
 * The PPM mode runs through imba_ppm_* (its camera paths sample only specular scattering). */
enum { IMBA_VCM_STATE_BYTES = 48 };
enum { IMBA_VCM_ALGO_VCM = 0, IMBA_VCM_ALGO_BPT = 1, IMBA_VCM_ALGO_LT = 3, IMBA_VCM_ALGO_PT = 4 };
typedef struct ImbaVcmFrame {
    const float* tri_normals;   /* device, 3 per triangle (as ImbaPtShading)                 */
    const int* tri_materials;
    int tri_count;
    const float* albedo;        /* device, 3 per material                                    */
    const float* emission;      /* device, 3 per material                                    */
    const float* emitter_area;  /* device, 1 per material                                    */
    const int* kinds;           /* device, 1 per material: 0 Lambertian, 1 mirror, 2 glass   */
    int material_count;
    const float* lights;        /* device, 12 per triangle light: p0, p1, p2, intensity      */
    int light_count;            /* >= 1                                                      */
    float raster_to_world[16];  /* imba_gen_camera_frame                                     */
    float world_to_raster[16];
    float cam_pos[3];
    float cam_dir[3];           /* the camera's unit forward direction                       */
    float image_plane_dist;
    int width, height, spp;
    int light_paths;            /* settings.light_path_count                                 */
    float radius;               /* the merge radius of this iteration                        */
    int max_path_len;
    int light_max_depth;
    int rr_start;               /* Russian roulette from this bounce on (0: the reference)   */
    int num_connections;        /* 1..8                                                      */
    int algo;                   /* IMBA_VCM_ALGO_*                                           */
} ImbaVcmFrame;

int imba_vcm_light(const ImbaVcmFrame* frame, uint32_t seed_base, int first, int count, ImbaRay* rays, void* states,
                   void* stream);
int imba_vcm_light_bounce(const ImbaVcmFrame* frame, const ImbaRay* rays, const ImbaHit* hits, void* states,
                          int capacity, const int* count_dev, void* vertices_a, void* vertices_b, ImbaRay* rays_out,
                          ImbaRay* shadow_rays, void* shadow_states, void* stream);
int imba_vcm_camera(const ImbaVcmFrame* frame, uint32_t seed_base, int first, int count, ImbaRay* rays, void* states,
                    void* stream);
int imba_vcm_camera_bounce(const ImbaVcmFrame* frame, const ImbaRay* rays, const ImbaHit* hits, void* states,
                           int capacity, const int* count_dev, const void* vertices_a, const void* vertices_b,
                           const int* vertex_count_dev, ImbaRay* rays_out, ImbaRay* shadow_rays, void* shadow_states,
                           int* shadow_count_dev, void* queries, float* image, void* stream);
int imba_vcm_merge_records(const ImbaHashGrid* grid, const void* vertices_a, const void* vertices_b, void* records,
                           void* stream);
int imba_vcm_merge(const ImbaVcmFrame* frame, const ImbaHashGrid* grid, const void* records, const void* queries,
                   int capacity, const int* count_dev, int k, float* image, void* stream);

/* ------------------------------ measurement ----------------------------------------------
 * (No counterpart in the reference: a probe for roofline reporting.)  Reads the device buffer
 * of `bytes` bytes `passes` times with 16-byte loads from every CU, each pass in a rotated
 * order so that a CU's vector L1 never serves a repeat, and returns the kernel's time in *ms
 * (so bytes * passes / ms is the read bandwidth of the cache level or HBM that holds a buffer
 * of that size).  sink: device scratch of at least 32 KiB * CUs bytes (less for small
 * buffers) that the kernel may write.  Synchronizes with the stream. */
int imba_read_bandwidth(const void* buffer, size_t bytes, int passes, void* sink, size_t sink_bytes, void* stream,
                        float* ms);

/* ------------------------------ synthetic workloads ---------------------------------------
 * Seeded generators for the measured configurations (host memory outputs). */
/* Random triangles in the unit cube: per triangle a centre ~ U[0,1)^3, then three vertices
 * centre + U[-0.01,0.01]^3, drawn in that order from the MWC64X generator of
 * render/random.h:15-47 seeded with `seed`.  vertices: 12*n floats, indices: 4*n. */
int imba_gen_random_triangles(uint64_t seed, int n, float* vertices, uint32_t* indices);
/* Icosphere: 5-level midpoint subdivision of an icosahedron projected to the sphere. */
int imba_gen_icosphere_counts(int level, int* vertex_count, int* tri_count);
int imba_gen_icosphere(int level, float radius, float cx, float cy, float cz, uint32_t material,
                       float* vertices, uint32_t* indices);
/* Pinhole camera rays (render/ray_gen/camera.h:19-79) for pixel centres, row-major,
 * tmin = 0, tmax = FLT_MAX. rays: host, w*h entries. */
int imba_gen_camera_rays(const float* pos3, const float* dir3, const float* up3, float fov_deg,
                         int width, int height, ImbaRay* rays);
/* The camera's raster -> world matrix (row major, applied to (raster y, raster x, 0, 1)). */
int imba_gen_camera_matrix(const float* pos3, const float* dir3, const float* up3, float fov_deg,
                           int width, int height, float* raster_to_world16);
/* The rest of PerspectiveCamera (camera.h:27-52, 66-69) for light tracing: world_to_raster
 * (row major, applied to (x, y, z, 1); raster x = t[1] / t[3], raster y = t[0] / t[3]), the
 * unit forward direction and the image plane distance width / (2 tan(fov / 2)). */
int imba_gen_camera_frame(const float* pos3, const float* dir3, const float* up3, float fov_deg,
                          int width, int height, float* world_to_raster16, float* forward3,
                          float* image_plane_dist);

#ifdef __cplusplus
}
#endif

#endif /* IMBA_HIP_H */
