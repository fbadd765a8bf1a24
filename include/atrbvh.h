/*
 * atrbvh.h -- C ABI of the MI355X (gfx950) BVH builder and traverser.
 *
 * Drop-in for the BVHRT-Core build path of leonardo-domingues/atrbvh.  Every
 * entry point names the reference interface it replaces (paths relative to
 * the reference repository root).  Conventions:
 *   - every function returns atr_status; atr_last_error() gives a
 *     thread-local message for the last failure on the calling thread;
 *   - all GPU work is ordered on the context's HIP stream and is
 *     asynchronous, except *_download, atr_sync and the calls documented as
 *     blocking;
 *   - handles are created and destroyed by the library; a context is used by
 *     one host thread at a time;
 *   - trees use the BVHRT SoA layout: internal nodes 0..n-2, leaves
 *     n-1..2n-2, root 0, parent(root) = -1, leaves have left = right = -1,
 *     internal nodes have data = -1, boxes are float4 with w = 0
 *     (BVHRT-Core/src/SoABVHTree.h:24-33).
 */
#ifndef ATRBVH_H
#define ATRBVH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum atr_status {
    ATR_OK = 0,
    ATR_ERR_INVALID_ARGUMENT = 1,
    ATR_ERR_OUT_OF_MEMORY = 2,
    ATR_ERR_DEVICE = 3,
    ATR_ERR_UNSUPPORTED = 4,
    ATR_ERR_INTERNAL = 5
} atr_status;

typedef struct atr_context atr_context;
typedef struct atr_scene atr_scene;
typedef struct atr_tree atr_tree;
typedef struct atr_collapsed atr_collapsed;

/* Host-side view of a tree (all arrays length 2n-1, boxes float4). */
typedef struct atr_tree_view {
    uint32_t num_triangles;
    int32_t root;
    int32_t* parent;
    int32_t* left;
    int32_t* right;
    int32_t* data;
    float* bbmin;
    float* bbmax;
    float* area;
} atr_tree_view;

/* Collapsed tree (BVHRT-Core/src/BVHCollapsedTree.h:9-44). Node arrays are
 * 2n-1 long, triangle_indices is n long and meaningful up to `used`. */
typedef struct atr_collapsed_view {
    uint32_t num_triangles;
    int32_t root;
    uint32_t used;
    int32_t* parent;
    int32_t* left;
    int32_t* right;
    int32_t* data;
    int32_t* triangle_count;
    float* bbmin;
    float* bbmax;
    int32_t* triangle_indices;
} atr_collapsed_view;

/* ------------------------------------------------------------------ misc */
const char* atr_last_error(void);
const char* atr_version(void);

/* Context: one HIP device, one stream, a caching device allocator.
 * Replaces the reference's implicit legacy stream and per-call cudaMalloc
 * (BVHRT-Core/src/TimeKernelExecution.h:14-32, LBVHBuilder.cpp:50-103). */
atr_status atr_context_create(int device, atr_context** out);
atr_status atr_context_destroy(atr_context* ctx);
atr_status atr_sync(atr_context* ctx);
/* Opaque HIP stream handle (hipStream_t) used by the context. */
void* atr_context_stream(atr_context* ctx);

/* Device memory helpers for callers without an allocator of their own.
 * Buffers come from the context's caching allocator. */
atr_status atr_malloc(atr_context* ctx, size_t bytes, void** dptr);
atr_status atr_free(atr_context* ctx, void* dptr);
atr_status atr_memcpy_h2d(atr_context* ctx, void* dst, const void* src, size_t bytes);
atr_status atr_memcpy_d2h(atr_context* ctx, void* dst, const void* src, size_t bytes);
atr_status atr_memset_d32(atr_context* ctx, void* dst, uint32_t value, size_t count);

/* Per-kernel timing with HIP events recorded on the context stream.  When
 * enabled every kernel launch is bracketed by two events; atr_profile_get
 * resolves them (blocking). */
atr_status atr_profile_enable(atr_context* ctx, int on);
atr_status atr_profile_reset(atr_context* ctx);
int atr_profile_count(atr_context* ctx);
atr_status atr_profile_get(atr_context* ctx, int i, const char** name, int* launches,
                           double* total_ms, double* bytes);

/* ----------------------------------------------------------------- scene */
/* Scene (BVHRT-Core/src/Scene.h:12-18) built like SceneWrapper
 * (BVHRT-Core/src/SceneWrapper.cpp:8-30): vertices are float4[3n] (xyz used).
 * The vertex array is COPIED to the device (SceneWrapper takes ownership of a
 * host array instead).  bbmin/bbmax NULL => the scene box is computed on the
 * device with SceneConverter's rule, including its running maximum starting
 * at FLT_MIN (gpu-ray-traversal/src/rt/SceneConverter.cpp:28-64).
 * 2 <= n <= 2^30 - 1 (more: ATR_ERR_UNSUPPORTED). */
atr_status atr_scene_create(atr_context* ctx, const float* vertices, uint32_t num_triangles,
                            const float* bbmin, const float* bbmax, atr_scene** out);
/* Same, from vertices already resident on the device (copied device-to-device).
 * Coordinates must be finite: host arrays are checked, device arrays are not
 * (a pass over them would cost a device-to-host round trip per call).  The
 * treelet passes are compiled on the premise that boxes, areas and SAH values
 * are finite; a scene with a non-finite coordinate has no defined tree. */
atr_status atr_scene_create_device(atr_context* ctx, const float* d_vertices,
                                   uint32_t num_triangles, const float* bbmin,
                                   const float* bbmax, atr_scene** out);
/* Scene box (blocking when it was computed on the device). */
atr_status atr_scene_bounds(atr_scene* scene, float bbmin[3], float bbmax[3]);
/* Device pointer of the scene's float4 vertex array (for refit/adjoint). */
atr_status atr_scene_vertices(atr_scene* scene, float** d_vertices);
atr_status atr_scene_destroy(atr_scene* scene);

/* --------------------------------------------------------------- builders */
/* LBVHBuilder(use64BitMortonCodes).BuildTree (BVHRT-Core/src/LBVHBuilder.h:31,
 * LBVHBuilder.cpp:21-130): Morton encode, stable key+index radix sort, radix
 * tree, bottom-up AABB refit.  morton_bits: 30 (lbvh) or 63 (lbvh64). */
atr_status atr_lbvh_build(atr_scene* scene, int morton_bits, atr_tree** out);

/* The triangle-splitting LBVH (BVHRT-Core/src/LBVH.h:36-70,96-155):
 * DeviceGenerateMortonCodes over primitive boxes (generateMortonCodesTSKernel /
 * generateMortonCodes64TSKernel, LBVH.cu:99-161), DeviceSort of the keys with
 * their routing indices, DeviceBuildTree with routing indices
 * (buildTreeTSKernel, LBVH.cu:320-428) and
 * DeviceCalculateInternalNodeBoundingBoxes (calculateNodeBoundingBoxesTSKernel,
 * LBVH.cu:572-669).  Input: num_boxes device boxes (float4 min / max, w
 * ignored), such as the fragments of split triangles, and for each the index
 * of the scene triangle it belongs to (device int32).  Keys come from the box
 * centroids in the scene's box; leaf i (sorted order) gets the box of its
 * routing index and that box's triangle as its data index, so several leaves
 * may name one triangle.  The collapse of such a tree writes each triangle
 * once per leaf list (TreeCollapser.cu:112-125).  The tree stays bound to
 * `scene` (its vertices serve collapse and traversal); atr_refit and
 * atr_tree_bind_scene reject it.  The arrays must stay valid until the work
 * queued on the context's stream has run. */
atr_status atr_lbvh_build_boxes(atr_scene* scene, const float* d_bbmin, const float* d_bbmax,
                                const int32_t* d_tri, uint32_t num_boxes, int morton_bits,
                                atr_tree** out);

/* TRBVHOptimizer(treeletSize, iterations).Optimize
 * (BVHRT-Core/src/TRBVHOptimizer.h:27, TRBVHOptimizer.cpp:21-128).
 * treelet_size must be 7. */
atr_status atr_trbvh_optimize(atr_tree* tree, int treelet_size, int iterations, float ci,
                              float ct);

/* AgglomerativeTreeletOptimizer(treeletSize, iterations).Optimize
 * (BVHRT-Core/src/AgglomerativeTreeletOptimizer.h:24,
 *  AgglomerativeTreeletOptimizer.cpp:24-110).  3 <= treelet_size <= 32. */
atr_status atr_atrbvh_optimize(atr_tree* tree, int treelet_size, int iterations, float ci,
                               float ct);

/* BVHTreeCollapser::Collapse (BVHRT-Core/src/BVHTreeCollapser.h:39,
 * BVHTreeCollapser.cpp:21-93).  *sah (nullable) = cost[root]/area[root]; when
 * sah is non-NULL the call blocks until it is known (as the reference does),
 * otherwise it is asynchronous and atr_collapsed_stats gives the SAH later.
 * The result is the reference's collapsed tree; when the tree is bound to a
 * scene, the collapsed tree keeps that scene's vertices for its traversal
 * layout (atr_collapsed_layout). */
atr_status atr_collapse(atr_tree* tree, float ci, float ct, atr_collapsed** out, float* sah);

/* The traversal layout atr_trace walks (64-byte node records and (v0, e1,
 * e2) triangle slots, built from the collapsed tree and its scene's vertices
 * on the GPU; the step CudaBVH takes on the host after the reference's
 * collapse, gpu-ray-traversal/src/rt/cuda/CudaBVH.cpp:469-574).  atr_trace
 * builds it on its first call; this builds it now (stream-ordered).  Fails
 * when the tree was not bound to a scene. */
atr_status atr_collapsed_layout(atr_collapsed* c);

/* ----------------------------------------------------------- stage calls */
/* The builders' stages one at a time, as the reference's low-level Device*
 * functions expose them.  Each queues its kernels on the context's stream and
 * takes its scratch from the context's caching allocator (the reference's
 * callers pass scratch in); results are bit-identical to the composite calls
 * above, which run the same kernels in the same order. */

/* DeviceBuildTree (BVHRT-Core/src/LBVH.h:80-95, buildTreeKernel /
 * buildTree64Kernel, LBVH.cu:208-316) and DeviceCalculateNodeBoundingBoxes
 * (LBVH.h:137-139, LBVH.cu:465-570) in one kernel: the radix tree over
 * num_triangles keys sorted ascending (d_sorted_keys: uint32 for morton_bits
 * 30, uint64 for 63, as atr_morton_encode + atr_sort_pairs give them), leaf i
 * holding triangle d_sort_indices[i] with that triangle's box, and every
 * internal node's box and area.  The tree is bound to `scene`, whose triangle
 * count must be num_triangles. */
atr_status atr_build_tree(atr_scene* scene, const void* d_sorted_keys,
                          const uint32_t* d_sort_indices, uint32_t num_triangles,
                          int morton_bits, atr_tree** out);

/* The triangle-splitting form: DeviceBuildTree with routing indices
 * (LBVH.h:108-125, buildTreeTSKernel, LBVH.cu:320-428) and
 * DeviceCalculateInternalNodeBoundingBoxes (LBVH.h:153-155, LBVH.cu:572-669):
 * leaf i gets box d_routing[i] of (d_bbmin, d_bbmax) and that box's triangle
 * d_tri[d_routing[i]] as its data index; num_boxes sorted keys.  Bound to
 * `scene` as atr_lbvh_build_boxes binds it. */
atr_status atr_build_tree_boxes(atr_scene* scene, const void* d_sorted_keys,
                                const uint32_t* d_routing, const float* d_bbmin,
                                const float* d_bbmax, const int32_t* d_tri, uint32_t num_boxes,
                                int morton_bits, atr_tree** out);

/* One TRBVH restructuring pass over every node with at least `gamma`
 * triangles: DeviceTreeletReestructureOptimizer (BVHRT-Core/src/TRBVH.h:42-46,
 * treeletReestructureKernel, TRBVH.cu:329-480).  TRBVHOptimizer runs it at
 * gamma = n, 2n, 4n, ... (TRBVHOptimizer.cpp:91,105).  treelet_size 7,
 * gamma >= treelet_size. */
atr_status atr_trbvh_pass(atr_tree* tree, int treelet_size, int gamma, float ci, float ct);

/* One ATRBVH restructuring pass: DeviceAgglomerativeTreeletOptimizer /
 * DeviceAgglomerativeSmallTreeletOptimizer (BVHRT-Core/src/AgglomerativeTreelet.h:39-75,
 * AgglomerativeTreelet.cu:214-595).  3 <= treelet_size <= 32,
 * gamma >= treelet_size. */
atr_status atr_atrbvh_pass(atr_tree* tree, int treelet_size, int gamma, float ci, float ct);

/* DeviceCopyTree + DeviceCalculateCosts (BVHRT-Core/src/TreeCollapser.h:19-42,
 * copyTreeKernel + calculateCostKernel, TreeCollapser.cu:13-93): the tree
 * copied into a new collapsed tree and every node's SAH cost and collapse
 * decision computed.  The result is unfinished until atr_collapse_lists runs
 * on it (every other call on it fails until then); `tree` must not change in
 * between. */
atr_status atr_collapse_costs(atr_tree* tree, float ci, float ct, atr_collapsed** out);

/* DeviceCollapseTree (TreeCollapser.h:60-62, collapseTreeKernel,
 * TreeCollapser.cu:173-237): the leaf lists of the nodes the cost stage chose,
 * which finishes `c`.  `tree` is the tree atr_collapse_costs read.  *sah
 * (nullable) as atr_collapse gives it. */
atr_status atr_collapse_lists(atr_tree* tree, atr_collapsed* c, float* sah);

/* SoABVHTree::SAH (BVHRT-Core/src/SoABVHTree.cpp:50-74), same summation order;
 * blocking (downloads the area array). */
atr_status atr_tree_sah(atr_tree* tree, float ci, float ct, float* out);

/* BVHTreeInstanceManager::DeviceToHostTree / FreeDeviceTree
 * (BVHRT-Core/src/BVHTreeInstanceManager.cpp:68-131).  The caller provides
 * host arrays in `host`; download is blocking. */
atr_status atr_tree_num_triangles(const atr_tree* tree, uint32_t* n);
atr_status atr_tree_download(atr_tree* tree, atr_tree_view* host);
atr_status atr_tree_device_view(atr_tree* tree, atr_tree_view* dev);
/* Upload (HostToDeviceTree, BVHTreeInstanceManager.cpp:56-66) after the
 * checks of atr_tree_view_check. */
atr_status atr_tree_upload(atr_tree* tree, const atr_tree_view* host);
/* Host-only validation of a tree view (no GPU needed): root 0; internal nodes
 * 0..n-2 with data -1; leaves n-1..2n-2 with left = right = -1 and data a
 * permutation of 0..n-1; every other node the child of exactly one internal
 * node; parent[] consistent with the links and -1 at the root; every node
 * reachable from the root.  The reference trusts its callers here
 * (BVHTreeInstanceManager.cpp:56-66); every kernel of this library walks the
 * links, so a tree that fails is rejected. */
atr_status atr_tree_view_check(const atr_tree_view* host);
atr_status atr_tree_destroy(atr_tree* tree);
/* Tree from a host view (HostToDeviceTree, BVHTreeInstanceManager.cpp:56-66). */
atr_status atr_tree_create_from_host(atr_context* ctx, const atr_tree_view* host,
                                     atr_tree** out);
/* Trees remember the scene they were built from; their collapsed trees use
 * its vertices for the traversal layout.  Uploaded trees are bound here.  The scene
 * must outlive the trees bound to it. */
atr_status atr_tree_bind_scene(atr_tree* tree, atr_scene* scene);

/* DeviceToHostCollapsedTree / FreeDeviceCollapsedTree
 * (BVHRT-Core/src/BVHTreeInstanceManager.cpp:196-270). Blocking. */
atr_status atr_collapsed_num_triangles(const atr_collapsed* c, uint32_t* n);
atr_status atr_collapsed_download(atr_collapsed* c, atr_collapsed_view* host);
/* SAH (cost[root]/area[root], BVHTreeCollapser.cpp:69-74), tree height and the
 * number of triangles placed in leaf lists.  Blocking. */
atr_status atr_collapsed_stats(atr_collapsed* c, float* sah, int* height, uint32_t* used);
atr_status atr_collapsed_destroy(atr_collapsed* c);

/* -------------------------------------------------------------- traversal */
/* Closest-hit (any_hit = 0) or any-hit traversal of a collapsed tree with
 * Moller-Trumbore intersection.  Replaces the `trace` kernel ABI
 * (gpu-ray-traversal/src/rt/kernels/CudaTracerKernels.hpp:73-86) launched by
 * CudaTracer::traceBatch (gpu-ray-traversal/src/rt/cuda/CudaTracer.cpp:80-158).
 * d_rays: device float4[2r] {ox,oy,oz,tmin, dx,dy,dz,tmax}
 *   (gpu-ray-traversal/src/rt/Util.hpp:62-71);
 * d_results: device int32[4r] {triangle id or -1, t bits, u bits, v bits};
 *   a miss stores t = tmax (RayResult, Util.hpp:77-87).
 * At most 2^31 rays per call (more: ATR_ERR_INVALID_ARGUMENT). */
atr_status atr_trace(atr_collapsed* c, const float* d_rays, uint32_t num_rays, int any_hit,
                     int32_t* d_results);

/* The reference tracer's BVH layout, built on the GPU from a collapsed tree:
 * CudaBVH::createCompact(const BVHCollapsedTree*, Scene*, nodeOffsetSizeDiv)
 * (gpu-ray-traversal/src/rt/cuda/CudaBVH.cpp:469-574, layout described in
 * CudaBVH.hpp:41-54) with Woop triangles (woopifyTri, :579-600), byte for
 * byte what that host DFS writes.  node_offset_size_div 1 = Compact (child
 * pointers in bytes), 16 = Compact2 (in 16-byte rows, what atr_trace_compact
 * and kepler_dynamic_fetch consume).  The arrays are device memory owned by
 * the collapsed tree (valid until it is destroyed or this is called again);
 * the call blocks while it reads back the sizes.  The tree must have been
 * bound to a scene. */
typedef struct atr_compact_layout {
    const float* nodes;         /* num_nodes x 16 words (4 x int4: child boxes, pointers) */
    uint32_t num_nodes;
    const float* tris;          /* num_rows x float4: 3 Woop rows per triangle,
                                   a row of 0x80000000 words ending each leaf */
    const int32_t* tri_indices; /* num_rows: triangle id on a triangle's first row */
    uint32_t num_rows;
    int node_offset_size_div;
} atr_compact_layout;
atr_status atr_collapsed_compact(atr_collapsed* c, int node_offset_size_div,
                                 atr_compact_layout* out);

/* The reference's kernel ABI for traversal (`trace`,
 * gpu-ray-traversal/src/rt/kernels/CudaTracerKernels.hpp:73-86) as a host
 * call on the context's stream, with kepler_dynamic_fetch semantics
 * (rt/kernels/kepler_dynamic_fetch.cu:61-320): Compact2 nodes in nodesA, Woop
 * rows in trisA, triIndices as written by atr_collapsed_compact; nodesB..D
 * and trisB..C are unused (as for the reference's Compact layouts).  rays:
 * float4[2 num_rays]; results: int4[num_rays] of which x = triangle id or -1
 * and y = t bits are written (the padding is left untouched, as the
 * reference's STORE_RESULT does).  A traversal deeper than 128 entries stops
 * that ray and raises the status that atr_sync reports. */
atr_status atr_trace_compact(atr_context* ctx, int num_rays, int any_hit, const float* rays,
                             int32_t* results, const float* nodesA, const float* nodesB,
                             const float* nodesC, const float* nodesD, const float* trisA,
                             const float* trisB, const float* trisC, const int32_t* triIndices);

/* Pinhole primary rays, row-major over a w x h grid: origin `eye`, direction
 * normalize(((2(x+.5)/w-1)*tan_half, (2(y+.5)/h-1)*tan_half, 1)) rotated into
 * the frame (right, up, forward) given by `basis` (9 floats, rows).
 * w x h at most 2^31 (more: ATR_ERR_INVALID_ARGUMENT); an empty grid writes
 * nothing. */
atr_status atr_rays_pinhole(atr_context* ctx, const float eye[3], const float basis[9],
                            float tan_half_fov, uint32_t width, uint32_t height, float tmax,
                            float* d_rays);
/* One cosine-weighted diffuse ray per primary ray, from its hit
 * (rayGenAOKernel semantics, gpu-ray-traversal/src/rt/ray/RayGenKernels.cu:114-221,
 * numSamples = 1): misses produce degenerate rays (tmax = -1). */
atr_status atr_rays_diffuse(atr_context* ctx, atr_scene* scene, const float* d_primary,
                            const int32_t* d_primary_results, uint32_t num_rays,
                            uint32_t seed, float tmax, float* d_rays);

/* RayBuffer::mortonSort (gpu-ray-traversal/src/rt/ray/RayBuffer.cpp:103-163;
 * findAABBKernel, genMortonKeysKernel, reorderRaysKernel,
 * rt/ray/RayBufferKernels.cu:64-196): sorts num_rays rays (float4[2] each)
 * by their 192-bit 6D Morton key -- origin in the rays' box (origins and end
 * points o + d*tmax; end points that are not finite are left out),
 * quantised to 2^24 steps, and (normalize(d) + 1)/2 quantised to 2^21 steps,
 * bit i of coordinate k at key bit k + 6i -- stably (equal keys keep slot
 * order; the reference sorts them on the host in no fixed order), all on the
 * GPU.  d_rays_out[i] is the ray of old slot p(i); d_slot_to_id[i] =
 * d_slot_to_id_in[p(i)] (p(i) when d_slot_to_id_in is NULL) and, when
 * d_id_to_slot is not NULL, d_id_to_slot[d_slot_to_id[i]] = i.  The keys use
 * IEEE division and square root (the reference builds with fast math).
 * d_rays_out must not overlap d_rays. */
atr_status atr_rays_morton_sort(atr_context* ctx, const float* d_rays, uint32_t num_rays,
                                const int32_t* d_slot_to_id_in, float* d_rays_out,
                                int32_t* d_slot_to_id, int32_t* d_id_to_slot);
/* The keys of atr_rays_morton_sort (genMortonKeysKernel): word w (key bits
 * 64w .. 64w+63) of ray i at d_keys[w * num_rays + i]; the rays' box (lo xyz,
 * hi xyz) in d_box (6 device floats). */
atr_status atr_rays_morton_keys(atr_context* ctx, const float* d_rays, uint32_t num_rays,
                                uint64_t* d_keys, float* d_box);
/* Results of sorted rays back in ray-id order: d_results[d_slot_to_id[s]] =
 * d_sorted_results[s] (int4 per ray; the lookup RayBuffer keeps in its
 * slot/id maps, RayBuffer.hpp). */
atr_status atr_rays_unsort_results(atr_context* ctx, const int32_t* d_sorted_results,
                                   const int32_t* d_slot_to_id, uint32_t num_rays,
                                   int32_t* d_results);

/* ------------------------------------------------------- low-level stages */
/* DeviceGenerateMortonCodes (BVHRT-Core/src/LBVH.h:20-50): keys (u32 for 30,
 * u64 for 63 bits) and indices (identity; may be NULL) on caller buffers.
 * n <= 2^30 - 1 (more: ATR_ERR_UNSUPPORTED). */
atr_status atr_morton_encode(atr_context* ctx, const float* d_vertices, uint32_t n,
                             const float bbmin[3], const float bbmax[3], int morton_bits,
                             void* d_keys, uint32_t* d_indices);
/* DeviceSort (BVHRT-Core/src/CubWrapper.h:18-34): stable ascending sort of
 * (key, value) pairs in place; key_bits 32 or 64 (key width in bits, all bits
 * sorted).  d_values NULL => values are treated as the identity and not
 * written. */
atr_status atr_sort_pairs(atr_context* ctx, int key_bits, void* d_keys, uint32_t* d_values,
                          uint32_t n);
/* atr_sort_pairs without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its scratch is a fixed 24 MiB device array per GPU, held until the process
 * ends; at most 1,048,576 pairs (larger n => ATR_ERR_INVALID_ARGUMENT).  Calls
 * that overlap on one GPU (two contexts or host threads on the same device)
 * race on that array.  Profiling events are not recorded for it. */
atr_status atr_sort_pairs_noalloc(atr_context* ctx, int key_bits, void* d_keys,
                                  uint32_t* d_values, uint32_t n);
/* Bottom-up refit of every box and area of a tree over the scene vertices
 * (calculateNodeBoundingBoxesKernel semantics, BVHRT-Core/src/LBVH.cu:465-570)
 * with the topology left unchanged. */
atr_status atr_refit(atr_tree* tree, atr_scene* scene);

/* ------------------------------------------------------- differentiation */
/* Adjoint of closest hits (no reference counterpart; the reference's tracer
 * is not differentiable).  For every ray r whose hit d_hits[4r] (as written
 * by atr_trace) is a triangle of `scene`, with upstream gradient
 * d_grad_tuv[3r..3r+2] = dL/d(t, u, v), adds dL/dv_i = -b_i M^{-T} g to the
 * gradient of the triangle's three vertices (b = (1-u-v, u, v),
 * M = [-d | v1-v0 | v2-v0]) with atomic adds into d_grad (float[3 x
 * num_grad_vertices], accumulated, not cleared).  d_index (uint32[3n],
 * nullable) maps triangle corners to mesh vertices (fixed topology); NULL =>
 * slot 3*triangle + corner.  Corners mapped at or beyond num_grad_vertices
 * are skipped. */
atr_status atr_hit_adjoint(atr_context* ctx, atr_scene* scene, const float* d_rays,
                           const int32_t* d_hits, uint32_t num_rays, const float* d_grad_tuv,
                           const uint32_t* d_index, uint32_t num_grad_vertices, float* d_grad);
/* Depth loss of the differentiable-refit workload: adds sum (t - t_target)^2
 * over rays hit in both d_hits and d_target_hits (and, when same_triangle is
 * nonzero, on the same triangle in both) to *d_loss (device double) and
 * writes d_grad_tuv = (2 (t - t_target), 0, 0) per such ray (zero elsewhere). */
atr_status atr_depth_loss(atr_context* ctx, const int32_t* d_hits, const int32_t* d_target_hits,
                          uint32_t num_rays, int same_triangle, float* d_grad_tuv,
                          double* d_loss);
/* Vertex step on an indexed mesh: every vertex of d_mesh_vertices
 * (float4[num_vertices]) moves by -lr * d_grad (float[3 num_vertices]),
 * shortened to length max_step when it is longer (max_step <= 0: no limit);
 * d_grad is then cleared (d_grad NULL => no step).  The scene's triangle soup
 * is then gathered from the mesh through d_index (uint32[3n]).  Follow with
 * atr_refit to update the tree.  Nothing here checks that the moved
 * coordinates stay finite: a step that sends one to inf or NaN leaves the
 * scene outside what the builders accept (see atr_scene_create_device). */
atr_status atr_mesh_step(atr_context* ctx, float* d_mesh_vertices, uint32_t num_vertices,
                         float* d_grad, float lr, float max_step, const uint32_t* d_index,
                         atr_scene* scene);

/* SAH of a tree into device memory, asynchronously, accumulated in double
 * precision by a parallel reduction (for monitoring such as a rebuild
 * trigger; atr_tree_sah gives the reference's exact fp32 summation order). */
atr_status atr_tree_sah_device(atr_tree* tree, float ci, float ct, float* d_out);

/* ---------------------------------------------------------- batched builds */
/* Many independent meshes built together.  No reference counterpart: it
 * replaces a host loop of LBVHBuilder::BuildTree (30-bit keys) +
 * AgglomerativeTreeletOptimizer::Optimize over meshes
 * (BVHRT-Core/src/LBVHBuilder.cpp:21-130,
 *  AgglomerativeTreeletOptimizer.cpp:24-110) with one launch sequence over
 * all meshes; every mesh's tree is bit-identical to the loop's.
 * vertices: float4[3T] mesh-major (host, or device for _create_device);
 * tri_offsets: host uint32[num_meshes + 1] prefix offsets, tri_offsets[0] = 0,
 * every mesh >= 2 triangles, T < 2^30; bboxes: host float[6 * num_meshes]
 * (min xyz, max xyz per mesh) or NULL => every build computes each mesh's box
 * on the device with SceneConverter's rule.  Coordinates must be finite: host
 * arrays are checked, device arrays are not. */
typedef struct atr_batch atr_batch;
atr_status atr_batch_create(atr_context* ctx, const float* vertices, const uint32_t* tri_offsets,
                            uint32_t num_meshes, const float* bboxes, atr_batch** out);
atr_status atr_batch_create_device(atr_context* ctx, const float* d_vertices,
                                   const uint32_t* tri_offsets, uint32_t num_meshes,
                                   const float* bboxes, atr_batch** out);
/* Device pointer of the batch's float4 vertex array (mesh-major); callers may
 * move vertices in place between builds, keeping every coordinate finite. */
atr_status atr_batch_vertices(atr_batch* b, float** d_vertices);
/* LBVH (30-bit) + ATRBVH(treelet, iterations) of every mesh, asynchronous.
 * use_graph != 0: the launch sequence is captured once into a HIP graph (per
 * parameter set) and replayed; it allocates nothing after the first call. */
atr_status atr_batch_build(atr_batch* b, int treelet_size, int iterations, float ci, float ct,
                           int use_graph);
/* Launch shape of batched builds (results do not depend on it): triangles per
 * thread of the encode kernel, 8 or 16 (default 16); keys per thread of a
 * radix-sort tile of 8192 keys, 8 (1024 threads) or 16 (512 threads)
 * (default 8); threads per workgroup of the treelet passes, 64, 128 or 256
 * (0 = the default, 128).
 * Changing it drops the captured graph.  Chosen per workload by
 * atrbvh_amd.autotune. */
typedef struct atr_batch_config {
    int32_t encode_items;
    int32_t sort_items;
    int32_t treelet_block;
} atr_batch_config;
atr_status atr_batch_set_config(atr_batch* b, const atr_batch_config* cfg);
atr_status atr_batch_get_config(const atr_batch* b, atr_batch_config* cfg);
/* SoABVHTree::SAH of every mesh (same summation order), into host out[num_meshes].
 * Blocking. */
atr_status atr_batch_sah(atr_batch* b, float ci, float ct, float* out);
/* Per-mesh SAH into device memory d_out[num_meshes], asynchronously, with
 * the double-precision parallel reduction of atr_tree_sah_device. */
atr_status atr_batch_sah_device(atr_batch* b, float ci, float ct, float* d_out);
/* Closest-hit traversal of every mesh on its binary tree: ray r (d_rays
 * float4[2R], R = rays_per_mesh * num_meshes) belongs to mesh
 * r / rays_per_mesh; d_hits int32[4R] as atr_trace writes them, with
 * triangle ids indexing the batch's vertex array.  Hits equal the per-mesh
 * collapsed traversal's except exact ties. */
atr_status atr_batch_trace(atr_batch* b, const float* d_rays, uint32_t rays_per_mesh,
                           int32_t* d_hits);
/* atr_hit_adjoint over the batch's triangle soup: d_grad is float[9T] (one
 * slot per triangle corner, accumulated). */
atr_status atr_batch_hit_adjoint(atr_batch* b, const float* d_rays, const int32_t* d_hits,
                                 uint32_t num_rays, const float* d_grad_tuv, float* d_grad);
/* Vertex step on the batch's soup: every corner moves by -lr * d_grad
 * (float[9T]), shortened to length max_step when longer (max_step <= 0: no
 * limit); d_grad is then cleared.  As for atr_mesh_step, the moved
 * coordinates must stay finite for later builds. */
atr_status atr_batch_step(atr_batch* b, float* d_grad, float lr, float max_step);
/* Bottom-up refit of every mesh's tree over the batch's current vertices. */
atr_status atr_batch_refit(atr_batch* b);
/* Rebuilds (LBVH 30-bit + ATRBVH) only the listed meshes from their current
 * vertices and puts their new trees in place; the other meshes keep theirs.
 * meshes: host uint32[count], distinct indices.  Allocates a temporary
 * sub-batch. */
atr_status atr_batch_rebuild(atr_batch* b, const uint32_t* meshes, uint32_t count,
                             int treelet_size, int iterations, float ci, float ct);
/* Mesh `mesh` as a standalone tree in the BVHRT layout (caller destroys it). */
atr_status atr_batch_extract(atr_batch* b, uint32_t mesh, atr_tree** out);
atr_status atr_batch_destroy(atr_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* ATRBVH_H */
