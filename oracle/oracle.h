/*
 * Serial CPU restatement of the BVHRT-Core build path and the closest-hit
 * traversal it feeds.  TEST INFRASTRUCTURE ONLY: this library is the checker
 * the HIP product path is compared against.  Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py may load it; the product never calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Build flags: -O2 -ffp-contract=off (no FMA contraction anywhere), so every
 * SAH / area expression is evaluated as separate IEEE fp32 multiply and add,
 * which is the contract the GPU kernels are compiled to as well.
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   morton        exhaustive 10-bit / sampled 21-bit bit-interleave closed form
 *   sort          numpy stable argsort of (key, index)
 *   karras        brute-force top-down highest-differing-bit split (index ties)
 *   refit         recursive union of leaf boxes
 *   sah           float sum in SoABVHTree::SAH order vs numpy float32 cumsum
 *   trbvh sched   compiled reference TRBVHScheduler.cpp (oracle/_ref) and
 *                 dependency-order invariants
 *   trbvh dp      exhaustive enumeration of all 10,395 topologies of 7 leaves
 *   atrbvh sched  pair coverage (each i>j pair exactly once)
 *   atrbvh        O(n^3) greedy agglomerative loop with the tie rule, on
 *                 synthetic treelets; tree invariants after optimisation
 *   tie order     closed-form tie rules of find_minimum, find_leaf_to_expand
 *                 and the DP partition choice, on planted ties
 *   composition   ATRBVH (n, 1) on n triangles = greedy clustering placed in
 *                 formTreelet's expansion order; TRBVH (7, 1) on 7 triangles =
 *                 exhaustive optimum of the DP objective; treelet counts per
 *                 gamma = n, 2n, 4n; planted mistakes (ORC_MUTATION) fail them
 *   parity unpinned: the slot order of rewritten nodes (formTreelet's two
 *                 children of an expanded node, TRBVH's left-then-right
 *                 hand-out per stack round) and the reference quirks that
 *                 feed later decisions (right-child cost, root nodesSah not
 *                 updated)
 *   collapse      recursive cost / leaf-set partition / SAH=cost/area
 *   trace         brute-force Moller-Trumbore over all triangles
 */
#ifndef ATRBVH_ORACLE_H
#define ATRBVH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* SoA tree, BVHRT index convention: internal nodes 0..n-2, leaves n-1..2n-2,
 * root 0, parent(root) = -1 (SoABVHTree.h:24-33, LBVH.cu:274-315).
 * bbmin/bbmax are float4 per node (w written as 0). */
typedef struct orc_tree {
    uint32_t n;
    int32_t root;
    int32_t* parent;
    int32_t* left;
    int32_t* right;
    int32_t* data;
    float* bbmin;
    float* bbmax;
    float* area;
} orc_tree;

/* BVHCollapsedTree (BVHCollapsedTree.h:9-44). */
typedef struct orc_collapsed {
    uint32_t n;
    int32_t root;
    int32_t* parent;
    int32_t* left;
    int32_t* right;
    int32_t* data;
    int32_t* count;
    float* bbmin;
    float* bbmax;
    int32_t* tri_index; /* n entries, only [0, used) meaningful */
    uint32_t used;      /* total triangles written to tri_index */
} orc_collapsed;

/* --- scene ------------------------------------------------------------ */
void orc_scene_bbox(const float* v4, uint32_t n, float bbmin[3], float bbmax[3]);

/* --- Morton codes ------------------------------------------------------ */
uint32_t orc_expand_bits10(uint32_t v);
uint64_t orc_expand_bits21(uint64_t v);
uint32_t orc_morton30(float x, float y, float z);
uint64_t orc_morton63(float x, float y, float z);
void orc_encode30(const float* v4, uint32_t n, const float bbmin[3], const float bbmax[3],
                  uint32_t* keys, uint32_t* idx);
void orc_encode63(const float* v4, uint32_t n, const float bbmin[3], const float bbmax[3],
                  uint64_t* keys, uint32_t* idx);

/* --- sort -------------------------------------------------------------- */
void orc_sort_pairs32(uint32_t* keys, uint32_t* vals, uint32_t n);
void orc_sort_pairs64(uint64_t* keys, uint32_t* vals, uint32_t n);

/* --- LBVH -------------------------------------------------------------- */
void orc_karras32(const uint32_t* keys, const uint32_t* idx, uint32_t n, orc_tree* t);
void orc_karras64(const uint64_t* keys, const uint32_t* idx, uint32_t n, orc_tree* t);
void orc_refit(const float* v4, orc_tree* t);
/* encode + sort + karras + refit; morton_bits 30 or 63 */
int orc_lbvh(const float* v4, uint32_t n, const float bbmin[3], const float bbmax[3],
             int morton_bits, orc_tree* t);
/* LBVH over primitive boxes (float4 min/max, m of them), each naming its
 * triangle in tri[]: the triangle-splitting variant (LBVH.cu:99-161,
 * 320-428, 572-669). */
int orc_lbvh_boxes(const float* bmin4, const float* bmax4, const int32_t* tri, uint32_t m,
                   const float bbmin[3], const float bbmax[3], int morton_bits, orc_tree* t);
float orc_sah(const orc_tree* t, float ci, float ct);

/* --- triangular matrix helpers (TriangularMatrix.cuh:34-52) ------------ */
int orc_lower_trm_row(int index);
int orc_lower_trm_col(int index);

/* --- schedulers -------------------------------------------------------- */
/* Writes rounds*warp ints into out (cap ints available); returns rounds or -1. */
int orc_trbvh_schedule(int treelet, int warp, int32_t* out, int cap);
/* Writes the packed pair schedule; returns its size or -1. */
int orc_atrbvh_schedule(int treelet, int warp, int32_t* out, int cap);

/* --- optimisers (in place) --------------------------------------------- */
int orc_trbvh_optimize(orc_tree* t, int treelet, int iterations, float ci, float ct);
int orc_atrbvh_optimize(orc_tree* t, int treelet, int iterations, float ci, float ct);

/* One bottom-up pass (kind 0 = TRBVH, 1 = ATRBVH) with an explicit treelet
 * threshold gamma; returns the number of treelets it processed, or -1. */
int orc_treelet_pass(orc_tree* t, int kind, int treelet, int gamma, float ci, float ct);

/* Single-treelet primitives exposed for the closed-form tests. */
float orc_trbvh_dp(int n, const float* leaf_cost, const float* subset_area, int treelet_tris,
                   float ci, float ct, float* costs_out, int32_t* masks_out);
/* The shuffle-reduction emulations used by the treelet passes, exposed for
 * their tests: findMinimumDistance over the first lastRow*(lastRow+1)/2
 * entries of D (AgglomerativeTreelet.cu:65-99) and findLeafToExpand over m
 * candidate areas (Treelet.cuh:23-40).  Return the winning index. */
int orc_find_minimum(const float* D, int lastRow);
int orc_find_leaf_to_expand(int m, const float* area);
int orc_atrbvh_cluster(int n, const float* boxes /* n*6 */, const float* leaf_sah, float ci,
                       int32_t* merges /* (n-1)*2 (row, col) */, float* areas_out /* n */);

/* --- collapse ---------------------------------------------------------- */
int orc_collapse(const orc_tree* t, float ci, float ct, orc_collapsed* out, float* sah);

/* --- traversal --------------------------------------------------------- */
/* rays: float4[2r] {ox,oy,oz,tmin, dx,dy,dz,tmax}; results int32[4r]
 * {tri id or -1, t bits, u bits, v bits}. */
void orc_trace(const orc_collapsed* c, const float* v4, const float* rays, uint32_t r,
               int any_hit, int32_t* results);
void orc_trace_brute(const float* v4, uint32_t n, const float* rays, uint32_t r,
                     int32_t* results);
void orc_ray_inverse(const float* d, float* inv);
int orc_ray_triangle(const float* v0, const float* v1, const float* v2, const float* o,
                     const float* d, float tmin, float tmax, float* t, float* u, float* v);

/* --- the reference tracer's layout (CudaBVH.cpp:469-600) ----------------- */
void orc_woopify(const float* v0, const float* v1, const float* v2, float out[12]);
/* createCompact with nodeOffsetSizeDiv div (1 = Compact, 16 = Compact2); NULL
 * outputs => sizes only.  nodes: 16 int32 per node; woop: 4 floats per row;
 * tri_index: one int32 per row. */
int orc_compact(const orc_collapsed* c, const float* v4, int div, int32_t* nodes,
                uint32_t* num_nodes, float* woop, int32_t* tri_index, uint32_t* num_rows);
/* kepler_dynamic_fetch trace semantics on a Compact2 layout. */
void orc_trace_compact(const int32_t* nodes, const float* woop, const int32_t* tri_index,
                       const float* rays, uint32_t r, int any_hit, int32_t* results);

/* --- ray generation ----------------------------------------------------- */
/* One diffuse ray per primary ray (rayGenAOKernel, RayGenKernels.cu:114-221,
 * numSamples = 1); prim float4[2r], res int32[4r] as orc_trace writes them. */
void orc_rays_diffuse(const float* v4, const float* prim, const int32_t* res, uint32_t r,
                      uint32_t seed, float tmax, float* out);

/* CPU baseline of the reference's CPU path (oracle/sbvh.cpp): SBVH build
 * (SplitBVHBuilder.cpp:52-476, "GPU" platform, leaves 1..8, alpha 1e-5) on
 * `threads` host threads, and a closest-hit tracer over it on `threads`
 * threads (BVH::traceRecursive semantics).  Results as orc_trace. */
typedef struct orc_sbvh_stats {
    uint32_t nodes;
    uint32_t tri_refs;
    uint32_t duplicates;
    int32_t depth;
    float sah;        /* top-down SAH, BVHNode.cpp:66-81 */
    float sah_bvhrt;  /* (1.2 inner areas + leaf areas x triangles) / root area */
} orc_sbvh_stats;
void* orc_sbvh_build(const float* v4, uint32_t n, int threads, orc_sbvh_stats* stats);
void orc_sbvh_trace(const void* bvh, const float* v4, const float* rays, uint32_t r, int threads,
                    int32_t* results);
void orc_sbvh_free(void* bvh);

/* --- ray coherence sort (RayBuffer.cpp:103-163, RayBufferKernels.cu:64-196)
 * rays: float[8] each; box: lo xyz, hi xyz; keys: word w of ray i at
 * keys[w*r + i] (three 64-bit words). */
void orc_ray_box(const float* rays, uint32_t r, float box[6]);
void orc_ray_keys(const float* rays, uint32_t r, const float box[6], uint64_t* keys);

#ifdef __cplusplus
}
#endif
#endif
