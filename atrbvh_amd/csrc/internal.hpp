// Internal declarations shared by the host API and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/atrbvh.h"

namespace atr {

// Default number of spin iterations of a bounded wait before it gives up and
// flags an error (a correct launch never gets close).
constexpr uint32_t kSpinLimit = 1u << 26;

// Number of b-sized groups that cover a items (a grid size).
inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// ------------------------------------------------------------------ errors
void set_error(const std::string& msg);
atr_status fail(atr_status code, const std::string& msg);

#define ATR_HIP(expr)                                                                    \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess)                                                            \
            return ::atr::fail(ATR_ERR_DEVICE, std::string(#expr) + ": " +              \
                                                   hipGetErrorString(_e));               \
    } while (0)

// Every C entry point is a function-try-block closed by ATR_CATCH_ALL, so
// that no C++ exception (a host allocation that fails, a container's length
// error) unwinds into a C, Go, Java or Python caller: it becomes
// ATR_ERR_OUT_OF_MEMORY (std::bad_alloc) or ATR_ERR_INTERNAL, with its
// message in atr_last_error().
atr_status fail_from_exception() noexcept;
#define ATR_CATCH_ALL \
    catch (...) { return ::atr::fail_from_exception(); }

#define ATR_TRY(expr)                                                                    \
    do {                                                                                 \
        atr_status _s = (expr);                                                          \
        if (_s != ATR_OK) return _s;                                                     \
    } while (0)

// Per-kernel profile record (events resolved lazily).
struct ProfileRecord {
    std::string name;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    int launches = 0;
    double total_ms = 0.0;
    double bytes = 0.0;
};

}  // namespace atr

struct atr_context {
    int device = 0;
    hipStream_t stream = nullptr;
    // caching allocator: freed blocks are kept per exact size
    std::multimap<size_t, void*> free_blocks;
    std::map<void*, size_t> live;
    // profiling
    bool profiling = false;
    std::vector<atr::ProfileRecord> prof;
    std::map<std::string, int> prof_index;
    std::vector<hipEvent_t> event_pool;
    // true while the stream is being captured into a HIP graph (no profiling
    // events are recorded then)
    bool capturing = false;
    // true inside an entry point that must allocate nothing (no profiling
    // events are created or recorded then either)
    bool no_alloc = false;
    // device status word: bit 0 = a bounded wait gave up, bit 1 = a tree was
    // deeper than a traversal stack, bit 2 = a traversal layout's counts were
    // inconsistent; cleared when atr_sync reports it
    uint32_t* d_status = nullptr;
    // iterations of a bounded wait before it gives up (ATRBVH_SPIN_LIMIT
    // overrides the default at context creation)
    uint32_t spin_limit = 0;
    // compute units of the device (0 until first queried)
    int cu_count = 0;
    // this device's copy of the no-allocation scratch array (noalloc.hip)
    void* noalloc_scratch = nullptr;
    std::string noalloc_error;  // why the lookup failed, if it did
    // pinned host staging for small readbacks, handed out in fixed slots
    void* h_pinned = nullptr;
    std::vector<int> pinned_free;
};

struct atr_scene {
    atr_context* ctx = nullptr;
    uint32_t n = 0;
    float4* verts = nullptr;  // 3n float4
    float* d_bbox = nullptr;  // 6 floats: min xyz, max xyz
    float h_bbox[6] = {0, 0, 0, 0, 0, 0};
    bool bbox_on_host = false;
};

struct atr_tree {
    atr_context* ctx = nullptr;
    atr_scene* scene = nullptr;  // scene the tree was built from (may be NULL)
    uint32_t n = 0;  // leaves (triangles)
    // Index of the first leaf node.  A single tree has internal nodes
    // 0..n-2 and first_leaf = n-1.  A forest of independent meshes (batched
    // builds) gives internal node slots 0..n-1 (the slot at each mesh's last
    // sorted position is unused) and first_leaf = n.
    uint32_t first_leaf = 0;
    int32_t* parent = nullptr;
    int32_t* left = nullptr;
    int32_t* right = nullptr;
    int32_t* data = nullptr;
    float4* bbmin = nullptr;
    float4* bbmax = nullptr;
    float* area = nullptr;
    // [min, max] leaf position (leaf node index - first_leaf) under each
    // node; a node whose span and index lie in one workgroup's block of leaves
    // (256, or 128 in the treelet passes) is handled by that workgroup only
    int2* span = nullptr;
    // built from primitive boxes (atr_lbvh_build_boxes): several leaves may
    // name one triangle, so collapsed leaf lists drop repeats, and a refit
    // from the scene's vertices would not give the leaves' boxes
    bool box_leaves = false;
    // A collapsed tree that reads this tree's parent / left / right / bbmin /
    // bbmax arrays in place (the collapse would copy them unchanged).  Before
    // any of them changes it gets its own copy; when this tree is destroyed
    // it takes them over.
    atr_collapsed* sharer = nullptr;
};

// Summary written by the collapse kernels into device memory.
struct CollapseSummary {
    float cost_root;
    float area_root;
    int32_t height;
    uint32_t used;
};

struct atr_collapsed {
    atr_context* ctx = nullptr;
    uint32_t n = 0;
    int32_t* parent = nullptr;
    int32_t* left = nullptr;
    int32_t* right = nullptr;
    int32_t* data = nullptr;
    int32_t* count = nullptr;
    float4* bbmin = nullptr;
    float4* bbmax = nullptr;
    int32_t* tri_index = nullptr;  // n
    // traversal layout
    float4* tnodes = nullptr;    // 4 float4 per internal node
    float4* ttris = nullptr;     // 3 float4 per slot (v0, e1, e2), 2n slots
    int32_t* ttri_id = nullptr;  // 2n
    bool has_layout = false;
    const float4* verts = nullptr;  // the bound scene's vertices (layout emission)
    // the reference tracer's Compact / Compact2 layout (atr_collapsed_compact)
    float4* cnodes = nullptr;       // 4 float4 per emitted node
    float4* cwoop = nullptr;        // Woop rows
    int32_t* crows_index = nullptr; // triangle id per Woop row
    uint32_t compact_nodes = 0;
    uint32_t compact_rows = 0;
    int compact_div = 0;
    CollapseSummary* d_summary = nullptr;
    CollapseSummary h_summary{};
    int pinned_slot = -1;
    bool summary_ready = false;
    hipEvent_t summary_event = nullptr;
    // between atr_collapse_costs and atr_collapse_lists: the tree the costs
    // were computed from and each internal node's list length
    const atr_tree* stage_tree = nullptr;
    uint32_t* stage_emit_count = nullptr;
    // the tree whose link and box arrays parent / left / right / bbmin /
    // bbmax point at (null when this collapsed tree owns them)
    atr_tree* links_of = nullptr;
};

// Many independent meshes built as one forest (batched builds).
struct atr_batch {
    atr_context* ctx = nullptr;
    uint32_t meshes = 0;
    uint32_t n = 0;                // total triangles
    std::vector<uint32_t> h_seg;   // meshes + 1 prefix offsets
    uint32_t* d_seg = nullptr;
    float4* verts = nullptr;       // 3n, owned
    float* d_bbox = nullptr;       // 6 per mesh
    bool bbox_given = false;
    std::vector<float> h_bbox;     // caller's boxes (6 per mesh) when given
    uint32_t* bbox_acc = nullptr;  // 6 per mesh
    atr_tree* forest = nullptr;    // first_leaf = n, 2n node slots
    // persistent workspace (every build, eager or graph, uses these)
    uint64_t* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    uint32_t* hist = nullptr;
    uint32_t hist_blocks = 0;
    int passes = 0;
    uint64_t* slots = nullptr;
    uint32_t* sort_offsets = nullptr;
    uint32_t* sort_status = nullptr;
    uint64_t* pass_slots = nullptr;  // arrival slots of the treelet passes (n)
    float* nsah = nullptr;
    int4* rec = nullptr;           // ATRBVH node records (n)
    uint32_t* counters = nullptr;
    float* d_sah = nullptr;        // per-mesh SAH
    bool built = false;
    // launch shape (atr_batch_config): triangles per encode thread, keys per
    // onesweep thread
    int encode_items = 16;
    int sort_items = 8;
    int treelet_block = 128;  // threads per workgroup of the treelet passes
    // captured build sequence
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int g_treelet = -1, g_iterations = -1;
    float g_ci = 0.0f, g_ct = 0.0f;
};

namespace atr {

// ------------------------------------------------------------- allocation
atr_status dalloc(atr_context* ctx, size_t bytes, void** p);
void dfree(atr_context* ctx, void* p);
template <typename T>
atr_status dalloc_t(atr_context* ctx, size_t count, T** p)
{
    return dalloc(ctx, count * sizeof(T), reinterpret_cast<void**>(p));
}

// ------------------------------------------------------------- launches
// Brackets a kernel launch with profiling events when enabled and checks the
// launch status.  `bytes` is the algorithmic byte count of the launch.
struct LaunchScope {
    atr_context* ctx;
    int rec;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    LaunchScope(atr_context* c, const char* name, double bytes);
    atr_status finish();
};

template <typename F>
atr_status launch(atr_context* ctx, const char* name, double bytes, F&& f)
{
    LaunchScope s(ctx, name, bytes);
    f();
    return s.finish();
}

// ------------------------------------------------------------- kernels
// LBVH (lbvh.hip)
atr_status run_scene_bbox(atr_context* ctx, const float4* verts, uint32_t n, float* d_bbox);
atr_status run_encode(atr_context* ctx, const float4* verts, uint32_t n, const float* d_bbox,
                      int bits, void* keys, uint32_t* idx, uint32_t* block_hist,
                      uint32_t hist_blocks, uint64_t* fill64, uint32_t fill64_count,
                      uint32_t* zero32, uint32_t zero32_count);
// Optional persistent sort workspace (digit offsets: passes*256 words;
// look-back status: sort_status_words(n, passes) words).
struct SortWork {
    uint32_t* offsets;
    uint32_t* status;
    // digit histograms when the caller has none (hist_words(n, passes) words;
    // may be NULL when block histograms are always given)
    uint32_t* hist = nullptr;
};
// Words of the per-block digit histograms of a sort of n keys.
size_t sort_hist_words(uint32_t n, int passes);
size_t sort_status_words(uint32_t n, int passes);
atr_status run_sort(atr_context* ctx, int key_bits, int passes, void* keys, uint32_t* vals,
                    uint32_t n, void* keys_alt, uint32_t* vals_alt, bool vals_identity,
                    const uint32_t* block_hist, uint32_t hist_blocks, const SortWork* work,
                    bool* result_in_alt, int tile_items = 8);
// Forests (batched builds): per-mesh boxes, mesh-tagged 30-bit keys, build.
int forest_passes(uint32_t meshes);
atr_status run_bbox_forest(atr_context* ctx, const float4* verts, const uint32_t* d_seg,
                           uint32_t meshes, uint32_t n, uint32_t* acc, float* d_bbox);
uint32_t encode_forest_blocks(uint32_t n, int items);
atr_status run_encode_forest(atr_context* ctx, const float4* verts, const uint32_t* d_seg,
                             uint32_t meshes, uint32_t n, const float* d_bbox, uint64_t* keys,
                             uint32_t* block_hist, uint64_t* fill64, uint32_t fill64_count,
                             int items);
atr_status run_build_forest(atr_context* ctx, const uint64_t* keys, const uint32_t* idx,
                            const float4* verts, atr_tree* t, const uint32_t* d_seg,
                            uint32_t meshes, uint64_t* slots);
uint32_t encode_blocks(uint32_t n);
atr_status run_build(atr_context* ctx, int bits, const void* keys, const uint32_t* idx,
                     uint32_t n, const float4* verts, atr_tree* t, uint64_t* slots);
// Builds over caller-given primitive boxes (split triangles): keys from the
// box centroids; leaves take the box and triangle of their routing index.
atr_status run_encode_boxes(atr_context* ctx, const float4* bmin, const float4* bmax, uint32_t n,
                            const float* d_bbox, int bits, void* keys, uint32_t* block_hist,
                            uint32_t hist_blocks, uint64_t* fill64, uint32_t fill64_count);
atr_status run_build_boxes(atr_context* ctx, int bits, const void* keys, const uint32_t* idx,
                           uint32_t n, const float4* bmin, const float4* bmax,
                           const int32_t* tri, atr_tree* t, uint64_t* slots);
atr_status run_refit(atr_context* ctx, atr_tree* t, const float4* verts, uint32_t* counters);
// Host: spans of a tree given as host arrays (post-order over the topology).
void host_spans(uint32_t n, const int32_t* left, const int32_t* right, int2* span);

// Batched builds (batch.hip)
atr_status run_forest_extract(atr_context* ctx, const atr_tree* f, uint32_t s, uint32_t nm,
                              atr_tree* t);
atr_status run_forest_splice(atr_context* ctx, const atr_tree* sub, const uint32_t* d_sseg,
                             uint32_t smeshes, const uint32_t* d_map, const uint32_t* d_fseg,
                             atr_tree* f);
atr_status run_trace_forest(atr_context* ctx, const atr_tree* f, const uint32_t* d_seg,
                            const float4* verts, const float4* rays, uint32_t r,
                            uint32_t rays_per_mesh, int32_t* results);
atr_status run_forest_sah(atr_context* ctx, const atr_tree* f, const uint32_t* d_seg,
                          uint32_t meshes, float ci, float ct, float* d_out);
atr_status run_sah_parallel(atr_context* ctx, const uint32_t* d_seg, uint32_t meshes,
                            uint32_t n_single, int first_leaf, const float* area, float ci,
                            float ct, float* d_out);

// Differentiable hits (adjoint.hip)
atr_status run_hit_adjoint(atr_context* ctx, const float4* v, uint32_t n_tris,
                           const float4* rays, const int4* hits, uint32_t R, const float* g_tuv,
                           const uint32_t* index, uint32_t n_grad, float* grad);
atr_status run_depth_loss(atr_context* ctx, const int4* hits, const int4* target, uint32_t R,
                          int same_tri, float* g_tuv, double* loss);
atr_status run_mesh_step(atr_context* ctx, float4* mesh, uint32_t V, float* grad, float lr,
                         float max_step, const uint32_t* index, uint32_t n_tris, float4* soup);

// Treelets (treelet.hip)
// `block`: threads per workgroup of the pass (64, 128 or 256; 128 unless a
// batch's launch shape says otherwise).
atr_status run_atrbvh_pass(atr_context* ctx, atr_tree* t, int treelet, int gamma, float ci,
                           float ct, float* nsah, uint64_t* slots, int4* rec, int block = 128);
// Node records {left, right, area(left), area(right)} that the treelet passes
// read while forming treelets (first_leaf entries); built before the first
// pass of every optimisation and kept current by the passes.
atr_status run_make_records(atr_context* ctx, atr_tree* t, int4* rec);
atr_status clear_slots(atr_context* ctx, uint64_t* slots, uint32_t count);
// slots: first_leaf 64-bit arrival slots of the passes (cleared by each pass);
// nsah: 2n-1 node SAH values.
atr_status run_trbvh_pass(atr_context* ctx, atr_tree* t, int gamma, float ci, float ct,
                          float* nsah, uint64_t* slots, int4* rec, const int32_t* d_schedule,
                          int rounds, int block = 128);

// Collapse + traversal layout (collapse.hip): the reference's collapsed tree,
// and, separately, the traversal layout built from it and the scene's
// vertices (on the first trace or on request).
atr_status run_collapse(atr_context* ctx, atr_tree* t, atr_collapsed* c, float ci, float ct);
atr_status run_collapse_costs(atr_context* ctx, atr_tree* t, atr_collapsed* c, float ci, float ct,
                              uint32_t* emit_count);
atr_status run_collapse_lists(atr_context* ctx, atr_tree* t, atr_collapsed* c,
                              const uint32_t* emit_count);
atr_status run_layout(atr_context* ctx, atr_collapsed* c, const float4* verts);
atr_status run_scan_u32(atr_context* ctx, const uint32_t* in, uint32_t n, uint32_t* out,
                        uint32_t* total);

// The reference tracer's layout and its traversal (compact.hip)
atr_status run_compact(atr_context* ctx, atr_collapsed* c, const float4* verts, int div);
atr_status run_trace_compact(atr_context* ctx, int num_rays, int any_hit, const float4* rays,
                             int2* results, const float4* nodes, const float4* tris,
                             const int32_t* tri_index);

// Traversal + ray generation (trace.hip)
atr_status run_trace(atr_context* ctx, atr_collapsed* c, const float4* rays, uint32_t r,
                     int any_hit, int32_t* results, int depth);
atr_status run_rays_pinhole(atr_context* ctx, const float* eye, const float* basis,
                            float tan_half, uint32_t w, uint32_t h, float tmax, float4* rays);
atr_status run_rays_diffuse(atr_context* ctx, const float4* verts, const float4* primary,
                            const int32_t* results, uint32_t r, uint32_t seed, float tmax,
                            float4* rays);

// Ray coherence sort (raysort.hip): 6D Morton keys (three 64-bit words per
// ray, word w of ray i at keys[w*r + i]) in the rays' box (6 floats).
atr_status run_ray_keys(atr_context* ctx, const float4* rays, uint32_t r, uint64_t* keys,
                        float* d_box);
atr_status run_ray_sort(atr_context* ctx, const float4* rays, uint32_t r,
                        const int32_t* in_slot_to_id, float4* out, int32_t* slot_to_id,
                        int32_t* id_to_slot);
atr_status run_unsort_results(atr_context* ctx, const int4* sorted, const int32_t* slot_to_id,
                              uint32_t r, int4* out);

// No-allocation entry points (noalloc.hip): the device address of this
// device's scratch array, looked up once per device.
atr_status noalloc_setup(int device, void** scratch);

// Host schedules (schedule.cpp)
int trbvh_schedule(int treelet, int warp, std::vector<int32_t>& out);

}  // namespace atr
