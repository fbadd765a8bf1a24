// Batched builds: per-mesh views of a forest built in one launch sequence.
//
// A forest holds M independent meshes in one mesh-major triangle array
// (seg[m] = first triangle of mesh m).  Internal node slot g < n belongs to
// the mesh whose sorted positions contain g; leaves start at first_leaf = n.
// Each mesh's tree, renumbered to its own BVHRT layout (internal 0..Nm-2,
// leaves Nm-1..2Nm-2, root 0), is exactly what a build of that mesh alone
// gives.
#include "device.hpp"
#include "internal.hpp"

namespace atr {

namespace {

// Renumbers one mesh of a forest into a standalone tree.
__global__ __launch_bounds__(256) void k_forest_extract(
    uint32_t s, uint32_t nm, int first_leaf, const int32_t* __restrict__ P,
    const int32_t* __restrict__ L, const int32_t* __restrict__ R, const int32_t* __restrict__ D,
    const float4* __restrict__ bmin, const float4* __restrict__ bmax,
    const float* __restrict__ A, const int2* __restrict__ SP, int32_t* parent, int32_t* left,
    int32_t* right, int32_t* data, float4* bbmin, float4* bbmax, float* area, int2* span)
{
    const uint32_t x = blockIdx.x * 256u + threadIdx.x;
    if (x >= 2u * nm - 1u) return;
    const int last = (int)nm - 1;
    const int g = (int)x < last ? (int)(s + x) : first_leaf + (int)s + ((int)x - last);
    auto conv = [&](int y) -> int32_t {
        if (y < 0) return -1;
        return y < first_leaf ? y - (int)s : last + (y - first_leaf - (int)s);
    };
    parent[x] = conv(P[g]);
    left[x] = conv(L[g]);
    right[x] = conv(R[g]);
    const int d = D[g];
    data[x] = d >= 0 ? d - (int)s : -1;
    bbmin[x] = bmin[g];
    bbmax[x] = bmax[g];
    area[x] = A[g];
    const int2 sp = SP[g];
    span[x] = make_int2(sp.x - (int)s, sp.y - (int)s);
}

// SoABVHTree::SAH (BVHRT-Core/src/SoABVHTree.cpp:50-74) of every mesh, one
// thread per mesh, summed serially in the reference's order: root, internal
// nodes 1..Nm-2, then the leaves.
__global__ __launch_bounds__(64) void k_forest_sah(const uint32_t* __restrict__ seg,
                                                   uint32_t meshes, int first_leaf,
                                                   const float* __restrict__ area, float ci,
                                                   float ct, float* out)
{
    const uint32_t m = blockIdx.x * 64u + threadIdx.x;
    if (m >= meshes) return;
    const uint32_t s = seg[m], e = seg[m + 1];
    const float total = area[s];
    // Loads go out eight at a time ahead of the additions, which stay serial
    // and in index order (the float sum is the reference's, bit for bit).
    float si = total;
    uint32_t g = s + 1;
    const uint32_t ge = e - 1;  // internal nodes s+1 .. e-2
    for (; g + 8 <= ge; g += 8) {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = area[g + k];
#pragma unroll
        for (int k = 0; k < 8; ++k) si += v[k];
    }
    for (; g < ge; ++g) si += area[g];
    float sl = 0.0f;
    const float* la = area + first_leaf;
    uint32_t p = s;
    for (; p + 8 <= e; p += 8) {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = la[p + k];
#pragma unroll
        for (int k = 0; k < 8; ++k) sl += v[k];
    }
    for (; p < e; ++p) sl += la[p];
    out[m] = (ci * si + ct * sl) / total;
}

// SAH of every mesh with double-precision accumulation, one block per mesh
// (for monitoring, e.g. a rebuild trigger; the summation order differs from
// SoABVHTree::SAH, so the last bits may too).  Mesh m: internal nodes
// [seg[m] + 1, seg[m] + Nm - 2] (root seg[m] excluded), leaves
// first_leaf + [seg[m], seg[m+1]).  A single tree of n leaves passes
// seg = NULL (one mesh [0, n)) and first_leaf = n - 1.
__global__ __launch_bounds__(256) void k_sah_parallel(const uint32_t* __restrict__ seg,
                                                      uint32_t n_single, int first_leaf,
                                                      const float* __restrict__ area, float ci,
                                                      float ct, float* out)
{
    __shared__ double s_i[4], s_l[4];
    const uint32_t m = blockIdx.x;
    const uint32_t s = seg ? seg[m] : 0u, e = seg ? seg[m + 1] : n_single;
    double si = 0.0, sl = 0.0;
    for (uint32_t g = s + 1 + threadIdx.x; g + 1 < e; g += 256u) si += area[g];
    for (uint32_t p = s + threadIdx.x; p < e; p += 256u) sl += area[first_leaf + p];
    for (int o = 32; o > 0; o >>= 1) {
        si += __shfl_xor(si, o, 64);
        sl += __shfl_xor(sl, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        s_i[threadIdx.x >> 6] = si;
        s_l[threadIdx.x >> 6] = sl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double root = area[s];
        const double ti = root + ((s_i[0] + s_i[1]) + (s_i[2] + s_i[3]));
        const double tl = (s_l[0] + s_l[1]) + (s_l[2] + s_l[3]);
        out[m] = (float)(((double)ci * ti + (double)ct * tl) / root);
    }
}

__device__ __forceinline__ uint32_t seg_of(const uint32_t* seg, uint32_t meshes, uint32_t i)
{
    uint32_t lo = 0, hi = meshes;  // seg[lo] <= i < seg[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (seg[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Copies the trees of a sub-batch (meshes rebuilt on their own) into the
// forest slots of the meshes they replace: sub-batch mesh j is forest mesh
// map[j]; every node, link, triangle index and span moves by the distance
// between the two meshes' first positions.
__global__ __launch_bounds__(256) void k_forest_splice(
    uint32_t sn, uint32_t smeshes, const uint32_t* __restrict__ sseg, int sfirst,
    const uint32_t* __restrict__ map, const uint32_t* __restrict__ fseg, int ffirst,
    const int32_t* __restrict__ P, const int32_t* __restrict__ L, const int32_t* __restrict__ R,
    const int32_t* __restrict__ D, const float4* __restrict__ bmin,
    const float4* __restrict__ bmax, const float* __restrict__ A, const int2* __restrict__ SP,
    int32_t* parent, int32_t* left, int32_t* right, int32_t* data, float4* bbmin, float4* bbmax,
    float* area, int2* span)
{
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    if (g >= 2u * sn) return;
    const bool leaf = (int)g >= sfirst;
    const uint32_t pos = leaf ? g - (uint32_t)sfirst : g;
    const uint32_t j = seg_of(sseg, smeshes, pos);
    const int delta = (int)fseg[map[j]] - (int)sseg[j];
    auto conv = [&](int y) -> int32_t {
        if (y < 0) return -1;
        return y < sfirst ? y + delta : ffirst + (y - sfirst) + delta;
    };
    const int dst = leaf ? ffirst + (int)pos + delta : (int)g + delta;
    parent[dst] = conv(P[g]);
    left[dst] = conv(L[g]);
    right[dst] = conv(R[g]);
    const int d = D[g];
    data[dst] = d >= 0 ? d + delta : -1;
    bbmin[dst] = bmin[g];
    bbmax[dst] = bmax[g];
    area[dst] = A[g];
    const int2 sp = SP[g];
    span[dst] = make_int2(sp.x + delta, sp.y + delta);
}

}  // namespace

atr_status run_forest_splice(atr_context* ctx, const atr_tree* sub, const uint32_t* d_sseg,
                             uint32_t smeshes, const uint32_t* d_map, const uint32_t* d_fseg,
                             atr_tree* f)
{
    return launch(ctx, "forest_splice", (double)sub->n * 2.0 * 104.0, [&] {
        k_forest_splice<<<div_up(2ull * sub->n, 256), 256, 0, ctx->stream>>>(
            sub->n, smeshes, d_sseg, (int)sub->first_leaf, d_map, d_fseg, (int)f->first_leaf,
            sub->parent, sub->left, sub->right, sub->data, sub->bbmin, sub->bbmax, sub->area,
            sub->span, f->parent, f->left, f->right, f->data, f->bbmin, f->bbmax, f->area,
            f->span);
    });
}

atr_status run_sah_parallel(atr_context* ctx, const uint32_t* d_seg, uint32_t meshes,
                            uint32_t n_single, int first_leaf, const float* area, float ci,
                            float ct, float* d_out)
{
    return launch(ctx, "sah_parallel", 0.0, [&] {
        k_sah_parallel<<<meshes, 256, 0, ctx->stream>>>(d_seg, n_single, first_leaf, area, ci, ct,
                                                         d_out);
    });
}

atr_status run_forest_extract(atr_context* ctx, const atr_tree* f, uint32_t s, uint32_t nm,
                              atr_tree* t)
{
    return launch(ctx, "forest_extract", (2.0 * nm - 1.0) * 104.0, [&] {
        k_forest_extract<<<div_up(2ull * nm - 1, 256), 256, 0, ctx->stream>>>(
            s, nm, (int)f->first_leaf, f->parent, f->left, f->right, f->data, f->bbmin, f->bbmax,
            f->area, f->span, t->parent, t->left, t->right, t->data, t->bbmin, t->bbmax, t->area,
            t->span);
    });
}

atr_status run_forest_sah(atr_context* ctx, const atr_tree* f, const uint32_t* d_seg,
                          uint32_t meshes, float ci, float ct, float* d_out)
{
    return launch(ctx, "forest_sah", (double)f->n * 8.0, [&] {
        k_forest_sah<<<div_up(meshes, 64), 64, 0, ctx->stream>>>(d_seg, meshes,
                                                                 (int)f->first_leaf, f->area, ci,
                                                                 ct, d_out);
    });
}

}  // namespace atr
