// SAH leaf collapse (BVHRT-Core/src/TreeCollapser.cu) and the traversal
// layout emitted from the collapsed tree on the GPU.
//
// Collapse semantics kept from the reference: a node P hands its children's
// triangles to leaf lists iff the climbs from both of its children reach it
// (every node strictly below on both sides is collapsible) and P itself is
// not collapsible (collapseTreeKernel :173-237).  The reference takes list
// offsets from an atomicAdd in arbitrary order; here they come from an
// exclusive scan over emitting parents in node order (left list first), so
// the output is deterministic.
//
// Traversal layout (this build's, consumed by trace.hip):
//   tnodes[4p+0] = (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y)
//   tnodes[4p+1] = (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y)
//   tnodes[4p+2] = (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)
//   tnodes[4p+3] = (ptr0, ptr1, 0, 0) as int: internal child index, or ~slot
// for a leaf (cf. the Compact layout, gpu-ray-traversal/src/rt/cuda/CudaBVH.hpp:41-54).
// Triangle slots hold (v0, v1-v0, v2-v0); slot i < n is the single-triangle
// leaf n-1+i, slots n.. hold the emitted lists in list order.  ttri_id[slot]
// is the original triangle index with bit 31 set on a leaf's last triangle.
#include "device.hpp"
#include "internal.hpp"

namespace atr {

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;

// ------------------------------------------------------------------ scan
__global__ __launch_bounds__(kScanBlock) void k_scan_tile_sums(const uint32_t* __restrict__ in,
                                                               uint32_t n,
                                                               uint32_t* __restrict__ sums)
{
    __shared__ uint32_t tmp[4];
    const uint32_t base = blockIdx.x * (uint32_t)kScanTile;
    uint32_t s = 0;
    for (int j = 0; j < kScanItems; ++j) {
        const uint32_t i = base + (uint32_t)j * kScanBlock + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint32_t total;
    block256_exclusive_sum(s, tmp, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_sums(uint32_t* sums, uint32_t count,
                                                          uint32_t* total_out)
{
    __shared__ uint32_t tmp[4];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < count; base += kScanBlock) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < count ? sums[i] : 0u;
        uint32_t total;
        const uint32_t ex = block256_exclusive_sum(v, tmp, &total);
        if (i < count) sums[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

// Tile-local exclusive scan in blocked order (thread t owns items
// t*16 .. t*16+15 of the tile) plus the tile's prefix.
__global__ __launch_bounds__(kScanBlock) void k_scan_apply(const uint32_t* __restrict__ in,
                                                           uint32_t n,
                                                           const uint32_t* __restrict__ sums,
                                                           uint32_t* __restrict__ out)
{
    __shared__ uint32_t tmp[4];
    const uint32_t base = blockIdx.x * (uint32_t)kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const uint32_t i = base + j;
        v[j] = i < n ? in[i] : 0u;
        s += v[j];
    }
    uint32_t run = sums[blockIdx.x] + block256_exclusive_sum(s, tmp, nullptr);
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const uint32_t i = base + j;
        if (i < n) out[i] = run;
        run += v[j];
    }
}

// ------------------------------------------------------------------ cost
// copyTreeKernel + calculateCostKernel (TreeCollapser.cu:13-93) in one
// bottom-up pass.  flags bit0 = collapsible, bit1 = "a climb leaves this
// node" (leaf, or both children deliver and the node is collapsible).  What a
// parent needs from a child -- triangle count, cost, flags, height -- is one
// 16-byte record: kept in LDS, in a slot named by the parent and the child's
// side, when the parent is block-local, and stored device-coherently in
// global memory when the parent's hand-over crosses workgroups (the other
// child's thread may then run on another XCD).  The copy of the tree's links
// and boxes (copy_links; skipped when the collapsed tree reads the tree's
// arrays in place) does not depend on the climb: thread i copies internal
// node i and leaf i up front, coalesced; the climb writes only the counts.  A
// node's area is computed from its box, as the reference does
// (TreeCollapser.cu:50-51), not taken from the tree's area array.
__global__ __launch_bounds__(256) void k_collapse_cost(
    uint32_t N, float ci, float ct, const int32_t* __restrict__ parent,
    const int32_t* __restrict__ left, const int32_t* __restrict__ right,
    const int32_t* __restrict__ data, const float4* __restrict__ bbmin,
    const float4* __restrict__ bbmax, const float* __restrict__ tarea, int32_t* c_parent,
    int32_t* c_left, int32_t* c_right, int32_t* c_data, int32_t* c_count, float4* c_bbmin,
    float4* c_bbmax, int4* st, uint32_t* emit_count, uint32_t* counters,
    const int2* __restrict__ span, CollapseSummary* summary, bool copy_links)
{
    __shared__ uint32_t s_cnt[256];
    // a block-local parent p's two child records: [2 (p & 255) + side]
    __shared__ int4 s_hand[512];
    s_cnt[threadIdx.x] = 0u;
    __syncthreads();
    const int blk = (int)blockIdx.x;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= N) return;
    if (i + 1 < N) {  // internal node i
        c_data[i] = -1;
        if (copy_links) {
            c_parent[i] = parent[i];
            c_left[i] = left[i];
            c_right[i] = right[i];
            c_bbmin[i] = bbmin[i];
            c_bbmax[i] = bbmax[i];
        }
    }
    int node = (int)(N - 1 + i);
    int cur = parent[node];
    // The next node's span, links, parent and box are loaded together as soon
    // as it is known (they do not change during the pass).
    int2 nsp = make_int2(0, 0);
    int nl = -1, nr = -1, npp = -1;
    float4 nmn = make_float4(0.0f, 0.0f, 0.0f, 0.0f), nmx = nmn;
    auto fetch = [&](int x) {
        nsp = span[x];
        nl = left[x];
        nr = right[x];
        npp = parent[x];
        nmn = bbmin[x];
        nmx = bbmax[x];
    };
    if (cur >= 0) fetch(cur);
    bool local = cur >= 0 && block_local(nsp, cur, blk);
    {
        const float4 mn = bbmin[node];
        const float4 mx = bbmax[node];
        if (copy_links) {
            c_parent[node] = cur;
            c_left[node] = left[node];
            c_right[node] = right[node];
            c_bbmin[node] = mn;
            c_bbmax[node] = mx;
        }
        c_data[node] = data[node];
        c_count[node] = 1;
    }
    // this thread's node's record, handed to the parent: in LDS (by parent
    // and side) when the parent is block-local, in global memory otherwise
    int4 mine = make_int4(1, __float_as_int(ct * box_area4(bbmin[node], bbmax[node])), 3, 0);
    if (cur >= 0) {
        if (local) s_hand[2 * (cur & 255) + (nl == node ? 0 : 1)] = mine;
        else st_coh(&st[node], mine);
    }
    while (cur >= 0) {
        // block-local nodes meet in LDS (workgroup scope), others globally
        uint32_t old;
        if (local) {
            old = arrive_local(&s_cnt[cur & 255]);
        } else {
            // a sibling already there (counter 1) is followed at once: its
            // state was stored before it arrived (see peek_crossing)
            old = ld<true>(&counters[cur]);
            if (old == 0u) old = arrive_crossing(&counters[cur]);
        }
        if (old == 0u) return;
        const bool is_left = nl == node;
        const int4 sib = local ? s_hand[2 * (cur & 255) + (is_left ? 1 : 0)]
                               : ld<true>(&st[is_left ? nr : nl]);
        const int4 a = is_left ? mine : sib;
        const int4 b = is_left ? sib : mine;
        const int cl = a.x, cr = b.x;
        const float kl = __int_as_float(a.y), kr = __int_as_float(b.y);
        const uint32_t fl = (uint32_t)a.z, fr = (uint32_t)b.z;
        const int count = cl + cr;
        const float ar = box_area4(nmn, nmx);
        float cc = ci * ar + kl + kr;
        const float asLeaf = ct * ar * (float)count;
        const bool collapsible = asLeaf < cc;
        if (collapsible) cc = asLeaf;
        const bool both = (fl & 2u) && (fr & 2u);
        const uint32_t f = (collapsible ? 1u : 0u) | ((both && collapsible) ? 2u : 0u);
        emit_count[cur] = (both && !collapsible)
                              ? (uint32_t)((cl > 1 ? cl : 0) + (cr > 1 ? cr : 0))
                              : 0u;
        const int h = 1 + max(a.w, b.w);
        const int pp = npp;
        c_count[cur] = count;
        if (pp < 0) {
            summary->cost_root = cc;
            summary->area_root = tarea[cur];
            summary->height = h;
        }
        node = cur;
        cur = pp;
        if (cur >= 0) fetch(cur);
        local = cur >= 0 && block_local(nsp, cur, blk);
        mine = make_int4(count, __float_as_int(cc), (int)f, h);
        if (cur >= 0) {
            if (local) s_hand[2 * (cur & 255) + (nl == node ? 0 : 1)] = mine;
            else st_coh(&st[node], mine);
        }
    }
}

// Right-first DFS of one child subtree into its list (writeSubtreeLeavesNonrecursive,
// TreeCollapser.cu:95-137, the pending stack living at the end of the list's
// own slots).  A triangle already in the list is not written again
// (:112-125); the list's count is then the number of distinct triangles.
// Trees built from triangles have one leaf per triangle, so the check is
// made only for trees built from primitive boxes (`dedup`).
__device__ void emit_list(uint32_t N, int child, int cnt, uint32_t off, bool dedup,
                          const int32_t* __restrict__ left, const int32_t* __restrict__ right,
                          const int32_t* __restrict__ data, int32_t* c_data, int32_t* c_count,
                          int32_t* tri_index)
{
    int32_t* buf = tri_index + off;
    buf[cnt - 1] = left[child];
    buf[cnt - 2] = right[child];
    int ss = 2;
    int pos = 0;
    while (ss > 0) {
        const int idx = buf[cnt - ss];
        --ss;
        if (idx >= (int)N - 1) {
            const int tri = data[idx];
            bool fresh = true;
            if (dedup)
                for (int q = 0; q < pos && fresh; ++q) fresh = buf[q] != tri;
            if (fresh) buf[pos++] = tri;
        } else {
            const int l = left[idx];
            const int r = right[idx];
            buf[cnt - 1 - ss++] = l;
            buf[cnt - 1 - ss++] = r;
        }
    }
    c_data[child] = (int32_t)off;
    c_count[child] = pos;
}

__global__ __launch_bounds__(256) void k_collapse_emit(
    uint32_t N, bool dedup, const uint32_t* __restrict__ emit_count,
    const uint32_t* __restrict__ emit_off, const int32_t* __restrict__ left,
    const int32_t* __restrict__ right, const int32_t* __restrict__ data, int32_t* c_data,
    int32_t* c_count, int32_t* tri_index)
{
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p + 1 >= N) return;
    if (emit_count[p] == 0u) return;
    uint32_t off = emit_off[p];
    const int l = left[p];
    const int r = right[p];
    const int cl = c_count[l];
    const int cr = c_count[r];
    if (cl > 1) {
        emit_list(N, l, cl, off, dedup, left, right, data, c_data, c_count, tri_index);
        off += (uint32_t)cl;
    }
    if (cr > 1) emit_list(N, r, cr, off, dedup, left, right, data, c_data, c_count, tri_index);
}

// Single-triangle slots (one per tree leaf, in leaf order).
__global__ __launch_bounds__(256) void k_leaf_slots(uint32_t N, const int32_t* __restrict__ data,
                                                    const float4* __restrict__ v, float4* ttris,
                                                    int32_t* ttri_id)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= N) return;
    const int tri = data[N - 1 + i];
    const float4 a = v[3ull * tri];
    const float4 b = v[3ull * tri + 1];
    const float4 c = v[3ull * tri + 2];
    ttris[3ull * i] = a;
    ttris[3ull * i + 1] = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, 0.0f);
    ttris[3ull * i + 2] = make_float4(c.x - a.x, c.y - a.y, c.z - a.z, 0.0f);
    ttri_id[i] = (int32_t)((uint32_t)tri | 0x80000000u);
}

// The slots of every collapsed list (one thread per internal node that holds
// a list: count triangles from tri_index at its data offset), end mark on
// the list's last triangle.
__global__ __launch_bounds__(256) void k_list_slots(uint32_t N, const int32_t* __restrict__ c_data,
                                                    const int32_t* __restrict__ c_count,
                                                    const int32_t* __restrict__ tri_index,
                                                    const float4* __restrict__ v, float4* ttris,
                                                    int32_t* ttri_id)
{
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p + 1 >= N) return;
    const int off = c_data[p];
    if (off < 0) return;
    const int cnt = c_count[p];
    for (int j = 0; j < cnt; ++j) {
        const uint32_t slot = N + (uint32_t)off + (uint32_t)j;
        const int tri = tri_index[off + j];
        const float4 a = v[3ull * tri];
        const float4 b = v[3ull * tri + 1];
        const float4 c = v[3ull * tri + 2];
        ttris[3ull * slot] = a;
        ttris[3ull * slot + 1] = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, 0.0f);
        ttris[3ull * slot + 2] = make_float4(c.x - a.x, c.y - a.y, c.z - a.z, 0.0f);
        ttri_id[slot] = j == cnt - 1 ? (int32_t)((uint32_t)tri | 0x80000000u) : tri;
    }
}

// One 64-byte traversal record per internal node.  Records of nodes inside
// collapsed subtrees are written too but never reached.
__global__ __launch_bounds__(256) void k_layout(uint32_t N, const int32_t* __restrict__ c_left,
                                                const int32_t* __restrict__ c_right,
                                                const int32_t* __restrict__ c_data,
                                                const float4* __restrict__ c_bbmin,
                                                const float4* __restrict__ c_bbmax,
                                                float4* tnodes)
{
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p + 1 >= N) return;
    const int l = c_left[p];
    const int r = c_right[p];
    const int last = (int)N - 1;
    auto ptr = [&](int ch) -> int {
        if (ch >= last) return ~(ch - last);
        const int d = c_data[ch];
        if (d >= 0) return ~((int)N + d);
        return ch;
    };
    const float4 a0 = c_bbmin[l], a1 = c_bbmax[l];
    const float4 b0 = c_bbmin[r], b1 = c_bbmax[r];
    tnodes[4ull * p + 0] = make_float4(a0.x, a1.x, a0.y, a1.y);
    tnodes[4ull * p + 1] = make_float4(b0.x, b1.x, b0.y, b1.y);
    tnodes[4ull * p + 2] = make_float4(a0.z, a1.z, b0.z, b1.z);
    tnodes[4ull * p + 3] = make_float4(__int_as_float(ptr(l)), __int_as_float(ptr(r)), 0.0f, 0.0f);
}

__global__ void k_fill_u32c(uint32_t* p, uint32_t v, uint64_t count)
{
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < count; i += 256ull * gridDim.x)
        p[i] = v;
}

}  // namespace

atr_status run_scan_u32(atr_context* ctx, const uint32_t* in, uint32_t n, uint32_t* out,
                        uint32_t* total)
{
    const uint32_t tiles = std::max(div_up(n, kScanTile), 1u);
    uint32_t* sums = nullptr;
    ATR_TRY(dalloc_t(ctx, tiles, &sums));
    ATR_TRY(launch(ctx, "scan_tile_sums", n * 4.0, [&] {
        k_scan_tile_sums<<<tiles, kScanBlock, 0, ctx->stream>>>(in, n, sums);
    }));
    ATR_TRY(launch(ctx, "scan_sums", tiles * 8.0, [&] {
        k_scan_sums<<<1, kScanBlock, 0, ctx->stream>>>(sums, tiles, total);
    }));
    ATR_TRY(launch(ctx, "scan_apply", n * 8.0, [&] {
        k_scan_apply<<<tiles, kScanBlock, 0, ctx->stream>>>(in, n, sums, out);
    }));
    dfree(ctx, sums);
    return ATR_OK;
}

// The copy and cost stage: emit_count[p] (N - 1 words) receives the number
// of list triangles internal node p emits (zero for most nodes).
atr_status run_collapse_costs(atr_context* ctx, atr_tree* t, atr_collapsed* c, float ci, float ct,
                              uint32_t* emit_count)
{
    const uint32_t N = t->n;
    const size_t nodes = 2ull * N - 1;
    int4* state = nullptr;
    uint32_t* counters = nullptr;
    atr_status s = dalloc_t(ctx, nodes, &state);
    if (s == ATR_OK) s = dalloc_t(ctx, N, &counters);
    // every internal node is finished exactly once, so emit_count[0, N-1)
    // is written by the cost kernel; only the arrival counters start at zero
    if (s == ATR_OK)
        s = launch(ctx, "collapse_zero", N * 4.0, [&] {
            k_fill_u32c<<<std::min<uint32_t>(div_up(N, 256), 2048u), 256, 0, ctx->stream>>>(
                counters, 0u, N);
        });
    if (s == ATR_OK)
        s = launch(ctx, "collapse_cost", (double)N * 2.0 * 100.0, [&] {
            k_collapse_cost<<<div_up(N, 256), 256, 0, ctx->stream>>>(
                N, ci, ct, t->parent, t->left, t->right, t->data, t->bbmin, t->bbmax, t->area,
                c->parent, c->left, c->right, c->data, c->count, c->bbmin, c->bbmax, state,
                emit_count, counters, t->span, c->d_summary, c->links_of != t);
        });
    dfree(ctx, state);
    dfree(ctx, counters);
    return s;
}

// The list stage: offsets by an exclusive scan of emit_count, then the lists.
atr_status run_collapse_lists(atr_context* ctx, atr_tree* t, atr_collapsed* c,
                              const uint32_t* emit_count)
{
    const uint32_t N = t->n;
    uint32_t* emit_off = nullptr;
    ATR_TRY(dalloc_t(ctx, N, &emit_off));
    atr_status s = run_scan_u32(ctx, emit_count, N - 1, emit_off, &c->d_summary->used);
    if (s == ATR_OK)
        s = launch(ctx, "collapse_emit", 0.0, [&] {
            k_collapse_emit<<<div_up(N, 256), 256, 0, ctx->stream>>>(
                N, t->box_leaves, emit_count, emit_off, t->left, t->right, t->data, c->data,
                c->count, c->tri_index);
        });
    dfree(ctx, emit_off);
    return s;
}

atr_status run_collapse(atr_context* ctx, atr_tree* t, atr_collapsed* c, float ci, float ct)
{
    uint32_t* emit_count = nullptr;
    ATR_TRY(dalloc_t(ctx, t->n, &emit_count));
    atr_status s = run_collapse_costs(ctx, t, c, ci, ct, emit_count);
    if (s == ATR_OK) s = run_collapse_lists(ctx, t, c, emit_count);
    dfree(ctx, emit_count);
    return s;
}

// The traversal layout of a collapsed tree: single-triangle slots, list
// slots and node records (built on the first trace, or on request).
atr_status run_layout(atr_context* ctx, atr_collapsed* c, const float4* verts)
{
    const uint32_t N = c->n;
    ATR_TRY(launch(ctx, "leaf_slots", N * 68.0, [&] {
        k_leaf_slots<<<div_up(N, 256), 256, 0, ctx->stream>>>(N, c->data, verts, c->ttris,
                                                               c->ttri_id);
    }));
    ATR_TRY(launch(ctx, "list_slots", 0.0, [&] {
        k_list_slots<<<div_up(N, 256), 256, 0, ctx->stream>>>(N, c->data, c->count, c->tri_index,
                                                               verts, c->ttris, c->ttri_id);
    }));
    return launch(ctx, "layout", (double)(N - 1) * 64.0 * 2.0, [&] {
        k_layout<<<div_up(N, 256), 256, 0, ctx->stream>>>(N, c->left, c->right, c->data,
                                                          c->bbmin, c->bbmax, c->tnodes);
    });
}

}  // namespace atr
