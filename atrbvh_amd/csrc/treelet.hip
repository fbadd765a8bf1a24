// Treelet restructuring on gfx950 (wave64): ATRBVH agglomerative clustering
// (BVHRT-Core/src/AgglomerativeTreelet.cu:392-554) and TRBVH subset DP
// (BVHRT-Core/src/TRBVH.cu:329-479).
//
// Both run the reference's bottom-up pass: one thread per leaf climbs through
// per-node arrival counters; the second arrival computes the node's triangle
// count and SAH and, when the count reaches gamma, the node roots a treelet.
// A wave of 64 lanes takes the voting treelet roots two at a time, one per
// 32-lane half.  Treelets voted in the same round are disjoint subtrees, so
// processing two concurrently equals the reference's one-at-a-time loop.
// Inside a half every cross-lane reduction resolves ties exactly as the
// reference's width-32 __shfl_down halving trees with strict comparisons do
// (findLeafToExpand Treelet.cuh:23-39, findMinimum AgglomerativeTreelet.cu:
// 65-81, reduceOptimal TRBVH.cu:23-35): findLeafToExpand and reduceOptimal
// run the same tree; findMinimum takes the minimum value with DPP row shifts
// and picks, from a ballot of the lanes holding it, the one with the
// smallest 5-bit-reversed lane index -- the lane that tree selects.
//
// ATRBVH forms treelets from 16-byte node records {left, right,
// area(left), area(right)} that k_make_records builds before the first pass
// and every treelet rewrite keeps current.
#include "device.hpp"
#include "internal.hpp"

namespace atr {

namespace {

// (row << 8 | col) of entry p of a lower-triangular matrix stored row by row
// without its diagonal: p = row*(row-1)/2 + col (TriangularMatrix.cuh:34-52),
// in 32-bit words so that wave-uniform lookups are scalar loads.
struct TrmTable32 {
    uint32_t rc[512];
};
constexpr TrmTable32 make_trm32()
{
    TrmTable32 t{};
    int p = 0;
    for (int r = 1; r < 32; ++r)
        for (int c = 0; c < r; ++c) t.rc[p++] = (uint32_t)((r << 8) | c);
    return t;
}
__constant__ TrmTable32 c_trm32 = make_trm32();

__device__ __forceinline__ int trm(int row, int col) { return ((row - 1) * row) / 2 + col; }

// Row of entry p of the same matrix in registers: the hardware square root is
// within one row of the answer for every p < 496 and the integer checks make
// it exact.
__device__ __forceinline__ int trm_row(int p)
{
    int r = (int)((1.0f + __builtin_amdgcn_sqrtf((float)(8 * p + 1))) * 0.5f);
    r -= (r * (r - 1)) / 2 > p ? 1 : 0;
    r += (r * (r + 1)) / 2 <= p ? 1 : 0;
    return r;
}

// Cross-lane helpers for two 32-lane halves working on different treelets.
// bcast_half: the value of lane k0 (lower half) or k1 (upper half) of the
// caller's own half; k0 and k1 are wave-uniform.
__device__ __forceinline__ bool upper_half() { return (__lane_id() & 32u) != 0u; }
__device__ __forceinline__ int bcast_half(int v, int k0, int k1, bool upper)
{
    const int a = __builtin_amdgcn_readlane(v, k0 & 31);
    const int b = __builtin_amdgcn_readlane(v, 32 + (k1 & 31));
    return upper ? b : a;
}
__device__ __forceinline__ float bcast_half(float v, int k0, int k1, bool upper)
{
    return __int_as_float(bcast_half(__float_as_int(v), k0, k1, upper));
}
// Minimum of v over lanes i, i + S of each 16-lane row (DPP row shift).  A
// lane whose partner lies outside the row gets an unspecified value; such
// lanes never feed lane 0 of the row, the only lane read afterwards.
template <int S>
__device__ __forceinline__ float min_row(float v)
{
    return fminf(v, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x100 | S, 0xF, 0xF,
                                                             false)));
}

// Among the lanes of a 32-bit lane mask, the one with the smallest 5-bit
// reversed index: lanes with bit 0 clear first, then bit 1, ... (0 if the
// mask is empty).
__device__ __forceinline__ int lane_min_br5(uint32_t m)
{
    const uint32_t pat[5] = {0x55555555u, 0x33333333u, 0x0F0F0F0Fu, 0x00FF00FFu, 0x0000FFFFu};
#pragma unroll
    for (int b = 0; b < 5; ++b) {
        const uint32_t t = m & pat[b];
        m = t ? t : m;
    }
    return m ? __builtin_ctz(m) : 0;
}

// Per-half LDS layout shared by both kernels' treelet formation.
struct HalfTreelet {
    int32_t* internal;  // [32]
    int32_t* leaf;      // [32]
    float* area;        // [32]
};

// formTreelet (Treelet.cuh:57-114) on one 32-lane half, from 16-byte node
// records {left, right, area(left), area(right)} that the passes keep next to
// the SoA arrays (one load per expansion instead of two dependent ones; the
// records hold the values the arrays hold, so the choices are the same).
// Lane 0 of the half performs the expansions; the largest stored area among
// the current treelet leaves (real leaves count as -FLT_MAX) is found with
// the reference's non-power-of-two halving tree.
template <bool COH>
__device__ __forceinline__ void form_treelet_rec(const HalfTreelet& h, int hl, int root, int n,
                                                 int first_leaf, const int4* rec)
{
    if (hl == 0) {
        const int4 q = ld<COH>(&rec[root]);
        h.internal[0] = root;
        h.leaf[0] = q.x;
        h.leaf[1] = q.y;
        h.area[0] = __int_as_float(q.z);
        h.area[1] = __int_as_float(q.w);
    }
    for (int it = 0; it < n - 2; ++it) {
        const int m = it + 2;
        int idx = hl;
        float a = -INFINITY;
        if (hl < m) {
            a = h.area[hl];
            if (h.leaf[hl] >= first_leaf) a = -FLT_MAX;
        }
        int cnt = m;
        int sh = m >> 1;
        while (cnt > 1) {
            const int oi = __shfl_down(idx, sh, 32);
            const float oa = __shfl_down(a, sh, 32);
            if (oa > a) {
                a = oa;
                idx = oi;
            }
            cnt = (cnt + 1) >> 1;
            sh = cnt >> 1;
        }
        if (hl == 0) {
            const int rep = h.leaf[idx];
            const int4 q = ld<COH>(&rec[rep]);
            h.internal[it + 1] = rep;
            h.leaf[idx] = q.x;
            h.leaf[it + 2] = q.y;
            h.area[idx] = __int_as_float(q.z);
            h.area[it + 2] = __int_as_float(q.w);
        }
    }
}

// formTreelet as above, with the records of the root's subtree down to depth
// kBfsDepth fetched first, one tree level per round with all of a level's
// loads in flight together: lane p < kBfsNodes of the half holds the node at heap
// position p (children of p at 2p+1 and 2p+2) and its record.  Expansions
// then take the chosen node's record from its lane; an expansion below that
// depth loads it.  The choices are identical to one load per expansion: the
// records hold the values the arrays hold, only the order of the loads
// changes (4 dependent rounds instead of n-1 for a balanced 9-leaf treelet).
constexpr int kBfsDepth = 3;
constexpr int kBfsNodes = (2 << kBfsDepth) - 1;
template <bool COH>
__device__ __forceinline__ void form_treelet_bfs(const HalfTreelet& h, int32_t* pos, int hl,
                                                 int root, int n, int first_leaf,
                                                 const int4* rec)
{
    const bool upper = (__lane_id() & 32u) != 0u;
    const int hbase = upper ? 32 : 0;
    int bn = hl == 0 ? root : -1;
    int4 bq = make_int4(-1, -1, 0, 0);
#pragma unroll
    for (int d = 0; d <= kBfsDepth; ++d) {
        const int lo = (1 << d) - 1, hi = (2 << d) - 1;
        if (hl >= lo && hl < hi && bn >= 0 && bn < first_leaf) bq = ld<COH>(&rec[bn]);
        if (d < kBfsDepth) {
            // lanes of depth d+1 take their node from the parent lane's record
            const int par = hbase + (((hl - 1) >> 1) & 31);
            const int cx = __shfl(bq.x, par, 64);
            const int cy = __shfl(bq.y, par, 64);
            if (hl >= hi && hl < (4 << d) - 1) bn = ((hl - 1) & 1) ? cy : cx;
        }
    }
    if (hl == 0) {
        h.internal[0] = root;
        h.leaf[0] = bq.x;
        h.leaf[1] = bq.y;
        h.area[0] = __int_as_float(bq.z);
        h.area[1] = __int_as_float(bq.w);
        pos[0] = 1 < kBfsNodes ? 1 : -1;
        pos[1] = 2 < kBfsNodes ? 2 : -1;
    }
    for (int it = 0; it < n - 2; ++it) {
        const int m = it + 2;
        int idx = hl;
        float a = -INFINITY;
        if (hl < m) {
            a = h.area[hl];
            if (h.leaf[hl] >= first_leaf) a = -FLT_MAX;
        }
        int cnt = m;
        int sh = m >> 1;
        while (cnt > 1) {
            const int oi = __shfl_down(idx, sh, 32);
            const float oa = __shfl_down(a, sh, 32);
            if (oa > a) {
                a = oa;
                idx = oi;
            }
            cnt = (cnt + 1) >> 1;
            sh = cnt >> 1;
        }
        int p = -1;
        int rep = -1;
        if (hl == 0) {
            rep = h.leaf[idx];
            p = pos[idx];
        }
        // the chosen node's heap lane in each half (wave-uniform)
        const int p0 = __builtin_amdgcn_readlane(p, 0);
        const int p1 = __builtin_amdgcn_readlane(p, 32);
        int4 q = make_int4(bcast_half(bq.x, p0, p1, upper), bcast_half(bq.y, p0, p1, upper),
                           bcast_half(bq.z, p0, p1, upper), bcast_half(bq.w, p0, p1, upper));
        if (hl == 0) {
            if (p < 0) q = ld<COH>(&rec[rep]);
            h.internal[it + 1] = rep;
            h.leaf[idx] = q.x;
            h.leaf[m] = q.y;
            h.area[idx] = __int_as_float(q.z);
            h.area[m] = __int_as_float(q.w);
            const int c = 2 * p + 1;
            pos[idx] = (p >= 0 && c < kBfsNodes) ? c : -1;
            pos[m] = (p >= 0 && c + 1 < kBfsNodes) ? c + 1 : -1;
        }
    }
}

// ------------------------------------------------------------------ ATRBVH
// Per-half LDS words: internal 32 + leaf 32 + area 32, then the slots' boxes
// as two float4 arrays {min.xyz, span.x} and {max.xyz, span.y} (32 each, so a
// slot's box and leaf span are two 16-byte reads, and lane k's own slot sits
// in a distinct bank group), then the distance matrix.  A half's words are a
// multiple of 4 so every half's float4 arrays stay 16-byte aligned.
constexpr int kAggFixedWords = 32 + 32 + 32 + 128 + 128;
constexpr int agg_half_words(int n) { return (kAggFixedWords + n * (n - 1) / 2 + 3) & ~3; }

// Returns the SAH the root holds afterwards (the agglomerated cost when the
// treelet is rewritten, else root_sah), the same in every lane of the half.
// COH: device-coherent loads (another XCD may have written the subtree in
// this launch); BFS: the breadth-first formation (fewer dependent rounds,
// more loads: for treelets whose latency is exposed) instead of one record
// load per expansion.  Stores are always device-coherent.
template <bool COH, bool BFS>
__device__ __forceinline__ float atrbvh_treelet(float* lds, int hl, int root, float root_sah, int n,
                                               int first_leaf, float ci, int32_t* parent,
                                               int32_t* left, int32_t* right, float4* bbmin,
                                               float4* bbmax, float* area, float* nsah,
                                               int2* span, int4* rec)
{
    HalfTreelet h{(int32_t*)lds, (int32_t*)lds + 32, lds + 64};
    float4* smn = (float4*)(lds + 96);   // {min.xyz, span.x as int bits}
    float4* smx = (float4*)(lds + 224);  // {max.xyz, span.y as int bits}
    float* D = lds + kAggFixedWords;
    if constexpr (BFS)  // the box words hold the heap positions until the boxes load
        form_treelet_bfs<COH>(h, (int32_t*)smn, hl, root, n, first_leaf, rec);
    else
        form_treelet_rec<COH>(h, hl, root, n, first_leaf, rec);

    // Each lane also carries the area of its current node (`my_area`), so
    // that every rewritten node's record gets its children's areas.
    int node = -1;
    float sah = 0.0f;
    float my_area = 0.0f;
    float rec_al = 0.0f, rec_ar = 0.0f;
    // lane k > 0 ends up holding internal node k's right child and SAH
    int out_right = -1;
    float out_sah = 0.0f;
    if (hl < n) {
        node = h.leaf[hl];
        my_area = h.area[hl];
        sah = ld<COH>(&nsah[node]);
        const float4 a = ld<COH>(&bbmin[node]);
        const float4 b = ld<COH>(&bbmax[node]);
        const int2 q = ld<COH>(&span[node]);
        smn[hl] = make_float4(a.x, a.y, a.z, __int_as_float(q.x));
        smx[hl] = make_float4(b.x, b.y, b.z, __int_as_float(q.y));
    }
    // calculateDistancesMatrix (AgglomerativeTreelet.cu:18-63): every pair i > j
    const int tsize = n * (n - 1) / 2;
    for (int p = hl; p < tsize; p += 32) {
        const int pr = trm_row(p);
        const int pc = p - (pr * (pr - 1)) / 2;
        const float4 ia = smn[pr], ib = smx[pr], ja = smn[pc], jb = smx[pc];
        D[p] = union_area(ia.x, ia.y, ia.z, ib.x, ib.y, ib.z, ja.x, ja.y, ja.z, jb.x, jb.y, jb.z);
    }

    const bool upper = upper_half();
    for (int L = n - 1; L > 0; --L) {
        // findMinimumDistance (:83-99)
        const int size = L * (L + 1) / 2;
        float md = FLT_MAX;
        int mi = 0;
        for (int j = hl; j < size; j += 32) {
            const float dd = D[j];
            if (dd < md) {
                md = dd;
                mi = j;
            }
        }
        // The reference's width-32 __shfl_down halving tree with strict
        // comparisons picks the smallest lane value and, among equal values,
        // the lane with the smallest 5-bit-reversed index (the last step
        // prefers even lanes, the one before lanes 0 mod 4, ...).  Here: the
        // minimum value of each half (order-free: DPP row shifts into each
        // 16-lane row's first lane, the half's two rows combined on the
        // scalar unit), a ballot of the lanes holding it, and the winning
        // lane picked from the ballot on the scalar unit.
        float mv = min_row<8>(md);
        mv = min_row<4>(mv);
        mv = min_row<2>(mv);
        mv = min_row<1>(mv);
        // distances are areas, never negative, so their bit patterns order as
        // unsigned integers and the two rows of each half combine on the
        // scalar unit
        const uint32_t q0 = (uint32_t)__builtin_amdgcn_readlane(__float_as_int(mv), 0);
        const uint32_t q1 = (uint32_t)__builtin_amdgcn_readlane(__float_as_int(mv), 16);
        const uint32_t q2 = (uint32_t)__builtin_amdgcn_readlane(__float_as_int(mv), 32);
        const uint32_t q3 = (uint32_t)__builtin_amdgcn_readlane(__float_as_int(mv), 48);
        const float m0 = __uint_as_float(q0 < q1 ? q0 : q1);
        const float m1 = __uint_as_float(q2 < q3 ? q2 : q3);
        const uint32_t w0 = (uint32_t)__ballot(md == m0);
        const uint32_t w1 = (uint32_t)(__ballot(md == m1) >> 32);
        // Each half's minimum index, as wave-uniform values (an idle upper
        // half yields a stale index, clamped into the table and never used).
        const int mi0 = __builtin_amdgcn_readlane(mi, lane_min_br5(w0));
        const int mi1 = __builtin_amdgcn_readlane(mi, 32 + lane_min_br5(w1));
        const int rc0 = (int)c_trm32.rc[(uint32_t)mi0 & 511u];
        const int rc1 = (int)c_trm32.rc[(uint32_t)mi1 & 511u];
        const int r0 = rc0 >> 8, c0 = rc0 & 255;
        const int r1 = rc1 >> 8, c1 = rc1 & 255;
        const int r = upper ? r1 : r0;
        const int c = upper ? c1 : c0;

        // updateState (:101-150)
        const float s = bcast_half(sah, r0, r1, upper) + bcast_half(sah, c0, c1, upper);
        const int li = bcast_half(node, r0, r1, upper);
        const int ri = bcast_half(node, c0, c1, upper);
        const float ali = bcast_half(my_area, r0, r1, upper);
        const float ari = bcast_half(my_area, c0, c1, upper);
        // the merged box (min / max of slots c and r, in that operand order)
        // and leaf span, in every lane
        const float4 ca = smn[c], cb = smx[c], ra = smn[r], rb = smx[r];
        const float amn0 = fminf(ca.x, ra.x), amn1 = fminf(ca.y, ra.y), amn2 = fminf(ca.z, ra.z);
        const float amx0 = fmaxf(cb.x, rb.x), amx1 = fmaxf(cb.y, rb.y), amx2 = fmaxf(cb.z, rb.z);
        const float4 mmn = make_float4(
            amn0, amn1, amn2, __int_as_float(min(__float_as_int(ra.w), __float_as_int(ca.w))));
        const float4 mmx = make_float4(
            amx0, amx1, amx2, __int_as_float(max(__float_as_int(rb.w), __float_as_int(cb.w))));
        if (hl == c) {
            node = h.internal[L - 1];
            smn[c] = mmn;
            smx[c] = mmx;
            const float ar = box_area(amn0, amn1, amn2, amx0, amx1, amx2);
            const float v = s + ci * ar;
            sah = v;
            my_area = ar;
        }
        const int lastIndex = bcast_half(node, L, L, upper);
        const float sahLast = bcast_half(sah, L, L, upper);
        const float areaLast = bcast_half(my_area, L, L, upper);
        const float4 la = smn[L], lb = smx[L];
        if (hl == r) {
            node = lastIndex;
            sah = sahLast;
            my_area = areaLast;
            smn[r] = la;
            smx[r] = lb;
        }
        if (hl == L) {
            node = li;
            out_right = ri;
            out_sah = s + ci * box_area(amn0, amn1, amn2, amx0, amx1, amx2);
            rec_al = ali;
            rec_ar = ari;
            smn[L] = mmn;
            smx[L] = mmx;
        }
        // updateDistancesMatrix (:152-188): the merged cluster (now slot c,
        // its box in registers) against every other remaining slot
        if (L > 1) {
            const bool cpy = hl < L && hl != r;
            float cp = 0.0f;
            if (cpy) cp = D[trm(L, hl)];
            if (cpy) D[trm(max(r, hl), min(r, hl))] = cp;
            if (hl < L && hl != c) {
                const float4 oa = smn[hl], ob = smx[hl];
                // union_area(box[max(c, hl)], box[min(c, hl)])
                D[trm(max(c, hl), min(c, hl))] =
                    c > hl ? union_area(amn0, amn1, amn2, amx0, amx1, amx2, oa.x, oa.y, oa.z, ob.x,
                                        ob.y, ob.z)
                           : union_area(oa.x, oa.y, oa.z, ob.x, ob.y, ob.z, amn0, amn1, amn2, amx0,
                                        amx1, amx2);
            }
        }
    }

    // updateTreelet (:190-212)
    const float cost = bcast_half(out_sah, 1, 1, upper);
    if (cost < root_sah) {
        if (hl >= 1 && hl < n) {
            const int ni = h.internal[hl - 1];
            const int rr = out_right;
            st_coh(&left[ni], node);
            st_coh(&right[ni], rr);
            st_coh(&parent[node], ni);
            st_coh(&parent[rr], ni);
            st_coh(&nsah[ni], out_sah);
            const float4 a = smn[hl], b = smx[hl];
            const float4 mn = make_float4(a.x, a.y, a.z, 0.0f);
            const float4 mx = make_float4(b.x, b.y, b.z, 0.0f);
            st_coh(&bbmin[ni], mn);
            st_coh(&bbmax[ni], mx);
            st_coh(&area[ni], box_area4(mn, mx));
            st_coh(&span[ni], make_int2(__float_as_int(a.w), __float_as_int(b.w)));
            st_coh(&rec[ni], make_int4(node, rr, __float_as_int(rec_al), __float_as_int(rec_ar)));
        }
    }
    return cost < root_sah ? cost : root_sah;
}

// ------------------------------------------------------------------ TRBVH
constexpr int kTrN = 7;
constexpr int kTrSub = 1 << kTrN;
// internal 8 + leaf 8 + area 8 + boxes 48 + leaf spans 16 + costs 128 +
// masks 128 + stack 16.  Subset boxes are not kept: the DP needs only their
// areas, and the n-2 boxes a rewrite emits are recomputed from the leaf boxes
// (min / max are exact, so any order gives the same box).
constexpr int kTrWords = 8 + 8 + 8 + 48 + 16 + kTrSub + kTrSub + 16;

// Box and leaf span of subset S of the treelet's leaves.
__device__ __forceinline__ void subset_box(const float* box, const int2* lsp, int S, float* mn,
                                           float* mx, int2& ss)
{
    mn[0] = mn[1] = mn[2] = FLT_MAX;
    mx[0] = mx[1] = mx[2] = -FLT_MAX;
    ss = make_int2(0x7fffffff, -1);
    for (int i = 0; i < kTrN; ++i) {
        if (S & (1 << i)) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                mn[k] = fminf(box[i * 6 + k], mn[k]);
                mx[k] = fmaxf(box[i * 6 + 3 + k], mx[k]);
            }
            ss.x = min(ss.x, lsp[i].x);
            ss.y = max(ss.y, lsp[i].y);
        }
    }
}

// Rewrites keep the node records the formation reads current: a rewritten
// node's record names its two new children and their areas (a subset's area
// for an internal child, the treelet leaf's area otherwise).
template <bool COH, bool BFS>
__device__ __forceinline__ void trbvh_treelet(float* lds, int hl, int root, float root_sah,
                                              int root_tris, int first_leaf, float ci, float ct,
                                              const int32_t* sched, int rounds, int32_t* parent,
                                              int32_t* left, int32_t* right, float4* bbmin,
                                              float4* bbmax, float* area, float* nsah,
                                              int2* span, int4* rec)
{
    HalfTreelet h{(int32_t*)lds, (int32_t*)lds + 8, lds + 16};
    float* box = lds + 24;
    int2* lsp = (int2*)(box + 48);
    float* cost = box + 48 + 16;
    int32_t* mask = (int32_t*)(cost + kTrSub);
    int32_t* stk = mask + kTrSub;
    // formTreelet writes at most index n-1 = 6 into the 8-entry arrays
    if constexpr (BFS)  // the box words hold the heap positions until the boxes load
        form_treelet_bfs<COH>(h, (int32_t*)box, hl, root, kTrN, first_leaf, rec);
    else
        form_treelet_rec<COH>(h, hl, root, kTrN, first_leaf, rec);

    if (hl < kTrN) {
        const int node = h.leaf[hl];
        const float4 a = ld<COH>(&bbmin[node]);
        const float4 b = ld<COH>(&bbmax[node]);
        box[hl * 6 + 0] = a.x;
        box[hl * 6 + 1] = a.y;
        box[hl * 6 + 2] = a.z;
        box[hl * 6 + 3] = b.x;
        box[hl * 6 + 4] = b.y;
        box[hl * 6 + 5] = b.z;
        lsp[hl] = ld<COH>(&span[node]);
    }
    // calculateSubsetSurfaceAreas (TRBVH.cu:37-98): lane covers 4 subsets
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int s = hl * 4 + j;
        float mn[3], mx[3];
        int2 ss;
        subset_box(box, lsp, s, mn, mx, ss);
        cost[s] = box_area(mn[0], mn[1], mn[2], mx[0], mx[1], mx[2]);
        mask[s] = 0;
    }
    // leaf costs (:438-443)
    if (hl < kTrN) cost[1 << hl] = ld<COH>(&nsah[h.leaf[hl]]);
    const float T = (float)root_tris;

    // processSchedule (:100-134): subsets of size 2..n-2, one per lane per round
    for (int rd = 0; rd < rounds; ++rd) {
        const int S = sched[rd * 32 + hl];
        if (S != 0) {
            float best = FLT_MAX;
            int bp = 0;
            const int delta = (S - 1) & S;
            int p = (-delta) & S;
            while (p != 0) {
                const float cc = cost[p] + cost[p ^ S];
                if (cc < best) {
                    best = cc;
                    bp = p;
                }
                p = (p - delta) & S;
            }
            const float a = cost[S];
            cost[S] = fminf(ci * a + best, ct * a * T);
            mask[S] = bp;
        }
    }

    // processSubsets, size n-1 (:139-187): 4 lanes per subset
    {
        const int k = hl >> 2;
        const int full = kTrSub - 1;
        const int S = full & ~(1 << (k < kTrN ? k : 0));
        const int leftMask = -(1 << max(k - 1, 0));
        float best = FLT_MAX;
        int bp = 0;
        if (hl < 4 * kTrN) {
            int base = (hl & 3) + 1;
            int p = (((base & leftMask) << 1) | (base & ~leftMask)) << 1;
            const int parts = (1 << (kTrN - 2)) - 1;
            for (int j = hl & 3; j < parts; j += 4) {
                const float cc = cost[p] + cost[p ^ S];
                if (cc < best) {
                    best = cc;
                    bp = p;
                }
                base += 4;
                p = (((base & leftMask) << 1) | (base & ~leftMask)) << 1;
            }
        }
#pragma unroll
        for (int sh = 2; sh > 0; sh >>= 1) {
            const float oc = __shfl_down(best, sh, 32);
            const int op = __shfl_down(bp, sh, 32);
            if (oc < best) {
                best = oc;
                bp = op;
            }
        }
        if (hl < 4 * kTrN && (hl & 3) == 0) {
            const float a = cost[S];
            cost[S] = fminf(ci * a + best, ct * a * T);
            mask[S] = bp;
        }
    }
    // size n (:192-219): 32 lanes
    {
        const int full = kTrSub - 1;
        float best = FLT_MAX;
        int bp = 0;
        int p = (hl + 1) * 2;
        const int parts = (1 << (kTrN - 1)) - 1;
        for (int j = hl; j < parts; j += 32) {
            const float cc = cost[p] + cost[p ^ full];
            if (cc < best) {
                best = cc;
                bp = p;
            }
            p += 64;
        }
#pragma unroll
        for (int sh = 16; sh > 0; sh >>= 1) {
            const float oc = __shfl_down(best, sh, 32);
            const int op = __shfl_down(bp, sh, 32);
            if (oc < best) {
                best = oc;
                bp = op;
            }
        }
        if (hl == 0) {
            const float a = cost[full];
            cost[full] = fminf(ci * a + best, ct * a * T);
            mask[full] = bp;
        }
    }

    // updateTreelet (:222-327).  Each stack round assigns treelet internal
    // nodes first to the internal left children in lane order, then to the
    // internal right children in lane order; new stack entries follow the
    // same order.  Ballot + prefix popcount make this deterministic.
    const int full = kTrSub - 1;
    if (cost[full] < root_sah) {
        const int halfShift = (__lane_id() >> 5) * 32;
        const uint32_t ltm = (1u << hl) - 1u;
        if (hl == 0) {
            stk[0] = h.internal[0];
            stk[8] = full;
        }
        int sp = 1;
        int next = 1;
        while (sp > 0) {
            const bool act = hl < sp;
            int sroot = 0, S = 0, P = 0, Q = 0;
            if (act) {
                sroot = stk[hl];
                S = stk[8 + hl];
                P = mask[S];
                Q = P ^ S;
            }
            const bool lint = act && __popc((uint32_t)P) > 1;
            const bool rint = act && __popc((uint32_t)Q) > 1;
            const uint32_t bl = (uint32_t)(__ballot(lint) >> halfShift);
            const uint32_t br = (uint32_t)(__ballot(rint) >> halfShift);
            const int nl = __popc(bl);
            const int nr = __popc(br);
            if (act) {
                int lchild, rchild;
                float lar, rar;
                if (lint) {
                    const int slot = __popc(bl & ltm);
                    lchild = h.internal[next + slot];
                    float mn[3], mx[3];
                    int2 ss;
                    subset_box(box, lsp, P, mn, mx, ss);
                    const float4 mn4 = make_float4(mn[0], mn[1], mn[2], 0.0f);
                    const float4 mx4 = make_float4(mx[0], mx[1], mx[2], 0.0f);
                    lar = box_area4(mn4, mx4);
                    st_coh(&bbmin[lchild], mn4);
                    st_coh(&bbmax[lchild], mx4);
                    st_coh(&area[lchild], lar);
                    st_coh(&nsah[lchild], cost[P]);
                    st_coh(&span[lchild], ss);
                    stk[slot] = lchild;
                    stk[8 + slot] = P;
                } else {
                    const int li = __ffs(P) - 1;
                    lchild = h.leaf[li];
                    lar = h.area[li];
                }
                st_coh(&left[sroot], lchild);
                st_coh(&parent[lchild], sroot);
                if (rint) {
                    const int slot = nl + __popc(br & ltm);
                    rchild = h.internal[next + slot];
                    float mn[3], mx[3];
                    int2 ss;
                    subset_box(box, lsp, Q, mn, mx, ss);
                    const float4 mn4 = make_float4(mn[0], mn[1], mn[2], 0.0f);
                    const float4 mx4 = make_float4(mx[0], mx[1], mx[2], 0.0f);
                    rar = box_area4(mn4, mx4);
                    st_coh(&bbmin[rchild], mn4);
                    st_coh(&bbmax[rchild], mx4);
                    st_coh(&area[rchild], rar);
                    // the reference stores costs[partition] for the right child too (TRBVH.cu:309)
                    st_coh(&nsah[rchild], cost[P]);
                    st_coh(&span[rchild], ss);
                    stk[slot] = rchild;
                    stk[8 + slot] = Q;
                } else {
                    const int ri = __ffs(Q) - 1;
                    rchild = h.leaf[ri];
                    rar = h.area[ri];
                }
                st_coh(&right[sroot], rchild);
                st_coh(&parent[rchild], sroot);
                st_coh(&rec[sroot], make_int4(lchild, rchild, __float_as_int(lar),
                                              __float_as_int(rar)));
            }
            next += nl + nr;
            sp = nl + nr;
        }
    }
}

// ------------------------------------------------------------------ driver
// Workgroups of BLOCK threads (one leaf each; 64, 128 or 256): smaller groups
// free their LDS sooner once most of their threads have stopped climbing; 128
// measured ~3% faster than 256 on the 2M-triangle build (64 was slower), and
// batched builds let the launch-shape tuner choose.  The trees do not depend
// on it: the block size only decides which nodes hand over through LDS.
template <int BLOCK>
struct PassShift;
template <>
struct PassShift<64> {
    static constexpr int value = 6;
};
template <>
struct PassShift<128> {
    static constexpr int value = 7;
};
template <>
struct PassShift<256> {
    static constexpr int value = 8;
};

// A node's count and SAH travel up with the climb: the first of its two
// child threads to arrive parks {count, SAH} of its child in the node's
// 64-bit slot (an LDS slot for block-local nodes, a global slot otherwise) and
// stops; the second takes them with the same exchange, so it needs no further
// load to finish the node.  Slots start empty (all ones: no count is
// 0xFFFFFFFF).
constexpr uint64_t kSlotEmpty = ~0ull;
// Roots of at least this many triangles form their treelets breadth-first
// (measured on the 2M-triangle build: 2% off the passes at 16384, less at
// 2048, a loss at 512 where the extra loads meet the full-occupancy phase).
constexpr int kBfsMinTris = 16384;
__device__ __forceinline__ uint64_t pack_count_sah(int count, float sah)
{
    return ((uint64_t)(uint32_t)count << 32) | __float_as_uint(sah);
}

template <bool AGG, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_treelet_pass(uint32_t n_leaves, int first_leaf, int n,
                                                      int gamma, float ci,
                                                      float ct, int32_t* parent, int32_t* left,
                                                      int32_t* right, float4* bbmin,
                                                      float4* bbmax, float* area,
                                                      float* nsah, uint64_t* slots,
                                                      int2* span, int4* rec,
                                                      const int32_t* sched, int rounds,
                                                      int half_words)
{
    constexpr int kShift = PassShift<BLOCK>::value;
    extern __shared__ __align__(16) float smem[];
    __shared__ uint64_t s_slot[BLOCK];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int half = lane >> 5;
    const int hl = lane & 31;
    float* lds = smem + ((tid >> 6) * 2 + half) * half_words;

    const uint32_t gid = blockIdx.x * (uint32_t)BLOCK + (uint32_t)tid;
    const int blk = (int)blockIdx.x;
    s_slot[tid] = kSlotEmpty;
    __syncthreads();
    // `node` is the node this thread holds (count `ncount`, SAH `nsah_v`);
    // `cur` its parent, the node the thread climbs to next.  A node's own
    // links, area, span and parent are unchanged from the start of the
    // launch until the node is processed (only treelets rooted at it or above
    // rewrite them), so the next node's fields are loaded one step ahead,
    // plainly, while the thread arrives at and works on the current one.
    int cur = -1;
    int node = -1;
    int ncount = 0;
    float nsah_v = 0.0f;
    int2 f_sp = make_int2(0, 0);
    int f_l = -1, f_up = -1;
    float f_ar = 0.0f;
    if (gid < n_leaves) {
        node = first_leaf + (int)gid;
        ncount = 1;
        nsah_v = ct * area[node];
        st_coh(&nsah[node], nsah_v);
        cur = parent[node];
        if (cur >= 0) {
            f_sp = span[cur];
            f_l = left[cur];
            f_ar = area[cur];
            f_up = parent[cur];
        }
    }
    // A thread whose node is a treelet root may hold it back for one step
    // (`held`): its arrival, count, SAH and locality are kept, and it joins
    // the next vote, where it can share the wave with another root.
    int tc = 0;
    float my_sah = 0.0f;
    bool local = true;
    int up = -1;
    bool held = false, was_held = false;
    while (__ballot(cur >= 0) != 0ull) {
        if (!held) {
            tc = 0;
            my_sah = 0.0f;
            local = true;
            up = -1;
        }
        if (cur >= 0 && !held) {
            // Nodes whose leaves all belong to this block meet in LDS; the rest
            // meet in global memory (cross-XCD hand-over, see device.hpp).
            const int2 sp = f_sp;
            const int l = f_l;
            const float ar = f_ar;
            up = f_up;
            local = block_local<kShift>(sp, cur, blk);
            const uint64_t mine = pack_count_sah(ncount, nsah_v);
            // the next node's fields, loaded while this arrival waits for the
            // thread's earlier stores to complete
            if (up >= 0) {
                f_sp = span[up];
                f_l = left[up];
                f_ar = area[up];
                f_up = parent[up];
            }
            uint64_t old;
            if (local) {
                wait_memory();
                old = __hip_atomic_exchange(&s_slot[cur & (BLOCK - 1)], mine, __ATOMIC_ACQ_REL,
                                            __HIP_MEMORY_SCOPE_WORKGROUP);
            } else {
                // a sibling already there is taken with one load (no store
                // wait, no exchange); otherwise arrive as usual
                old = peek_crossing(&slots[cur]);
                if (old == kSlotEmpty) old = exchange_crossing(&slots[cur], mine);
            }
            if (old == kSlotEmpty) {
                cur = -1;
            } else {
                const int oc = (int)(uint32_t)(old >> 32);
                const float os = __uint_as_float((uint32_t)old);
                // SAH of the node: ci * area + SAH(left) + SAH(right), in this order
                const bool is_left = l == node;
                const float sl = is_left ? nsah_v : os;
                const float sr = is_left ? os : nsah_v;
                tc = ncount + oc;
                my_sah = ci * ar + sl + sr;
                st_coh(&nsah[cur], my_sah);
            }
        }
        uint64_t vote = __ballot(tc >= gamma);
        held = false;
        // Two roots share a wave (one per half).  With an odd count, the
        // highest-lane root waits one step if other lanes are still climbing
        // (they may bring a partner), it has not waited before and its
        // subtree is small (the long chain near the root is not delayed).
        // Only the TRBVH pass does this: it shortens that pass by ~2%, while
        // the cheaper agglomerative treelets measured ~3% slower with it.
        if (!AGG && (__popcll(vote) & 1) != 0) {
            const int last = 63 - (int)__clzll((unsigned long long)vote);
            const bool last_waited = __shfl((int)was_held, last, 64) != 0;
            const int last_tc = __shfl(tc, last, 64);
            if (!last_waited && last_tc < 2048 && __ballot(cur >= 0 && tc < gamma) != 0ull) {
                vote &= ~(1ull << last);
                if (lane == last) held = true;
            }
        }
        was_held = held;
        while (vote != 0ull) {
            const int r0 = __ffsll((unsigned long long)vote) - 1;
            vote &= vote - 1ull;
            int r1 = -1;
            if (vote != 0ull) {
                r1 = __ffsll((unsigned long long)vote) - 1;
                vote &= vote - 1ull;
            }
            const int my_r = half ? r1 : r0;
            const int src = my_r >= 0 ? my_r : r0;
            const int root = __shfl(cur, src, 64);
            const float root_sah = __shfl(my_sah, src, 64);
            const int root_tris = __shfl(tc, src, 64);
            const bool root_local = __shfl((int)local, src, 64) != 0;
            // The two halves share one instruction stream: when either root
            // is crossing, both treelets use the device-coherent loads (which
            // are also correct for a block-local treelet), so the wave runs
            // one flavour instead of both in turn.
            const bool wave_local = __ballot(my_r >= 0 && !root_local) == 0ull;
            // Large subtrees are few and their treelets sit on the pass's
            // longest chain (towards the root), where latency, not issue
            // rate, decides: they form breadth-first.
            const bool wave_big = __ballot(my_r >= 0 && root_tris >= kBfsMinTris) != 0ull;
            // the root's SAH after the treelet (what its parent adds up)
            float new_sah = root_sah;
            if (my_r >= 0) {
                if (AGG) {
                    if (wave_big)
                        new_sah = atrbvh_treelet<true, true>(lds, hl, root, root_sah, n,
                                                                   first_leaf, ci, parent, left,
                                                                   right, bbmin, bbmax, area,
                                                                   nsah, span, rec);
                    else if (wave_local)
                        new_sah = atrbvh_treelet<false, false>(lds, hl, root, root_sah, n,
                                                                     first_leaf, ci, parent, left,
                                                                     right, bbmin, bbmax, area,
                                                                     nsah, span, rec);
                    else
                        new_sah = atrbvh_treelet<true, false>(lds, hl, root, root_sah, n,
                                                                    first_leaf, ci, parent, left,
                                                                    right, bbmin, bbmax, area,
                                                                    nsah, span, rec);
                } else {
                    if (wave_big)
                        trbvh_treelet<true, true>(lds, hl, root, root_sah, root_tris, first_leaf,
                                                  ci, ct, sched, rounds, parent, left, right,
                                                  bbmin, bbmax, area, nsah, span, rec);
                    else if (wave_local)
                        trbvh_treelet<false, false>(lds, hl, root, root_sah, root_tris,
                                                    first_leaf, ci, ct, sched, rounds, parent,
                                                    left, right, bbmin, bbmax, area, nsah, span,
                                                    rec);
                    else
                        trbvh_treelet<true, false>(lds, hl, root, root_sah, root_tris, first_leaf,
                                                   ci, ct, sched, rounds, parent, left, right,
                                                   bbmin, bbmax, area, nsah, span, rec);
                }
            }
            const float s0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(new_sah), 0));
            const float s1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(new_sah), 32));
            if (lane == r0) my_sah = s0;
            if (lane == r1) my_sah = s1;
        }
        if (cur >= 0 && !held) {
            node = cur;
            ncount = tc;
            nsah_v = my_sah;
            cur = up;
        }
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t count)
{
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < count; i += 256ull * gridDim.x)
        p[i] = v;
}


// Node records {left, right, area(left), area(right)} of every internal node
// from the SoA arrays.  Internal slots whose links are not node indices (the
// spare slot of each mesh in a batch's forest) get {-1, -1, 0, 0}; no treelet
// reaches them.
__global__ __launch_bounds__(256) void k_make_records(uint32_t first_leaf, uint32_t nodes,
                                                      const int32_t* __restrict__ left,
                                                      const int32_t* __restrict__ right,
                                                      const float* __restrict__ area,
                                                      int4* __restrict__ rec)
{
    const uint32_t x = blockIdx.x * 256u + threadIdx.x;
    if (x >= first_leaf) return;
    const int l = left[x];
    const int r = right[x];
    int4 q = make_int4(-1, -1, 0, 0);
    if ((uint32_t)l < nodes && (uint32_t)r < nodes)
        q = make_int4(l, r, __float_as_int(area[l]), __float_as_int(area[r]));
    rec[x] = q;
}

}  // namespace

atr_status run_make_records(atr_context* ctx, atr_tree* t, int4* rec)
{
    const uint32_t fl = t->first_leaf;
    return launch(ctx, "make_records", (double)fl * 32.0, [&] {
        k_make_records<<<std::max(div_up(fl, 256), 1u), 256, 0, ctx->stream>>>(
            fl, fl + t->n, t->left, t->right, t->area, rec);
    });
}

template <bool AGG, int BLOCK>
static void launch_pass(atr_context* ctx, atr_tree* t, int n, int gamma, float ci, float ct,
                        float* nsah, uint64_t* slots, int4* rec, const int32_t* sched,
                        int rounds, int half_words)
{
    const uint32_t N = t->n;
    const size_t lds = (size_t)half_words * (BLOCK / 32) * sizeof(float);
    k_treelet_pass<AGG, BLOCK><<<div_up(N, BLOCK), BLOCK, lds, ctx->stream>>>(
        N, (int)t->first_leaf, n, gamma, ci, ct, t->parent, t->left, t->right, t->bbmin,
        t->bbmax, t->area, nsah, slots, t->span, rec, sched, rounds, half_words);
}

template <bool AGG>
static atr_status launch_pass_block(atr_context* ctx, const char* name, int block, atr_tree* t,
                                    int n, int gamma, float ci, float ct, float* nsah,
                                    uint64_t* slots, int4* rec, const int32_t* sched, int rounds,
                                    int half_words)
{
    if (block != 64 && block != 128 && block != 256)
        return fail(ATR_ERR_INVALID_ARGUMENT, "treelet pass block must be 64, 128 or 256");
    return launch(ctx, name, 0.0, [&] {
        if (block == 64)
            launch_pass<AGG, 64>(ctx, t, n, gamma, ci, ct, nsah, slots, rec, sched, rounds,
                                 half_words);
        else if (block == 128)
            launch_pass<AGG, 128>(ctx, t, n, gamma, ci, ct, nsah, slots, rec, sched, rounds,
                                  half_words);
        else
            launch_pass<AGG, 256>(ctx, t, n, gamma, ci, ct, nsah, slots, rec, sched, rounds,
                                  half_words);
    });
}

// Every arrival slot empty (all ones) before a pass.
atr_status clear_slots(atr_context* ctx, uint64_t* slots, uint32_t count)
{
    return launch(ctx, "clear_slots", count * 8.0, [&] {
        k_fill_u32<<<std::min<uint32_t>(std::max(div_up(2ull * count, 256), 1u), 2048u), 256, 0,
                     ctx->stream>>>((uint32_t*)slots, 0xFFFFFFFFu, 2ull * count);
    });
}

atr_status run_atrbvh_pass(atr_context* ctx, atr_tree* t, int treelet, int gamma, float ci,
                           float ct, float* nsah, uint64_t* slots, int4* rec, int block)
{
    ATR_TRY(clear_slots(ctx, slots, t->first_leaf));
    const int half_words = agg_half_words(treelet);
    return launch_pass_block<true>(ctx, "atrbvh_pass", block, t, treelet, gamma, ci, ct, nsah,
                                   slots, rec, nullptr, 0, half_words);
}

atr_status run_trbvh_pass(atr_context* ctx, atr_tree* t, int gamma, float ci, float ct,
                          float* nsah, uint64_t* slots, int4* rec, const int32_t* d_schedule,
                          int rounds, int block)
{
    ATR_TRY(clear_slots(ctx, slots, t->first_leaf));
    return launch_pass_block<false>(ctx, "trbvh_pass", block, t, kTrN, gamma, ci, ct, nsah,
                                    slots, rec, d_schedule, rounds, kTrWords);
}

}  // namespace atr
