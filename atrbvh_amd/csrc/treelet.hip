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

// formTreelet (Treelet.cuh:57-114) on one 32-lane half.  Lane 0 of the half
// performs the expansions; the largest stored area among the current treelet
// leaves (real leaves count as -FLT_MAX) is found with the reference's
// non-power-of-two halving tree.
template <bool COH>
__device__ __forceinline__ void form_treelet(const HalfTreelet& h, int hl, int root, int n,
                                             int first_leaf, const int32_t* left,
                                             const int32_t* right, const float* area)
{
    if (hl == 0) {
        const int l = ld<COH>(&left[root]);
        const int r = ld<COH>(&right[root]);
        h.internal[0] = root;
        h.leaf[0] = l;
        h.leaf[1] = r;
        h.area[0] = ld<COH>(&area[l]);
        h.area[1] = ld<COH>(&area[r]);
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
            const int l = ld<COH>(&left[rep]);
            const int r = ld<COH>(&right[rep]);
            h.internal[it + 1] = rep;
            h.leaf[idx] = l;
            h.leaf[it + 2] = r;
            h.area[idx] = ld<COH>(&area[l]);
            h.area[it + 2] = ld<COH>(&area[r]);
        }
    }
}

// formTreelet as above, reading each expanded node's children and their
// areas with one 16-byte load of its record {left, right, area(left),
// area(right)} (kept by the ATRBVH pass next to the SoA arrays) instead of
// two dependent loads.  The choices are identical: the record holds the
// values the arrays hold.
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

// ------------------------------------------------------------------ ATRBVH
// Per-half LDS words: internal 32 + leaf 32 + area 32 + boxes 32*6 + spans
// 32*2 + matrix.
constexpr int kAggFixedWords = 32 + 32 + 32 + 192 + 64;

template <bool COH>
__device__ __forceinline__ void atrbvh_treelet(float* lds, int hl, int root, float root_sah, int n,
                                               int first_leaf, float ci, int32_t* parent,
                                               int32_t* left, int32_t* right, float4* bbmin,
                                               float4* bbmax, float* area, float* nsah,
                                               int2* span, int4* rec)
{
    HalfTreelet h{(int32_t*)lds, (int32_t*)lds + 32, lds + 64};
    float* box = lds + 96;
    int2* sp = (int2*)(lds + 288);
    float* D = lds + kAggFixedWords;
    form_treelet_rec<COH>(h, hl, root, n, first_leaf, rec);

    // Each lane also carries the area of its current node (`my_area`), so
    // that every rewritten node's record gets its children's areas.
    int node = -1;
    float sah = 0.0f;
    float my_area = 0.0f;
    float rec_al = 0.0f, rec_ar = 0.0f;
    if (hl < n) {
        node = h.leaf[hl];
        my_area = h.area[hl];
        sah = ld<COH>(&nsah[node]);
        const float4 a = ld<COH>(&bbmin[node]);
        const float4 b = ld<COH>(&bbmax[node]);
        box[hl * 6 + 0] = a.x;
        box[hl * 6 + 1] = a.y;
        box[hl * 6 + 2] = a.z;
        box[hl * 6 + 3] = b.x;
        box[hl * 6 + 4] = b.y;
        box[hl * 6 + 5] = b.z;
        sp[hl] = ld<COH>(&span[node]);
    }
    // calculateDistancesMatrix (AgglomerativeTreelet.cu:18-63): every pair i > j
    const int tsize = n * (n - 1) / 2;
    for (int p = hl; p < tsize; p += 32) {
        const int pr = trm_row(p);
        const float* bi = box + pr * 6;
        const float* bj = box + (p - (pr * (pr - 1)) / 2) * 6;
        D[p] = union_area(bi[0], bi[1], bi[2], bi[3], bi[4], bi[5], bj[0], bj[1], bj[2], bj[3],
                          bj[4], bj[5]);
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
        // minimum value of each half (order-free: xor-16 swizzle, then DPP
        // row shifts into lane 0), a ballot of the lanes holding it, and the
        // winning lane picked from the ballot on the scalar unit.
        float mv = fminf(md, __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(md), 0x401F)));
        mv = min_row<8>(mv);
        mv = min_row<4>(mv);
        mv = min_row<2>(mv);
        mv = min_row<1>(mv);
        const float m0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mv), 0));
        const float m1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mv), 32));
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
        float amn[3], amx[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            amn[k] = fminf(box[c * 6 + k], box[r * 6 + k]);
            amx[k] = fmaxf(box[c * 6 + 3 + k], box[r * 6 + 3 + k]);
        }
        const int2 spr = sp[r], spc = sp[c];
        const int2 spm = make_int2(min(spr.x, spc.x), max(spr.y, spc.y));
        if (hl == c) {
            node = h.internal[L - 1];
            sp[c] = spm;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                box[c * 6 + k] = amn[k];
                box[c * 6 + 3 + k] = amx[k];
            }
            const float ar = box_area(amn[0], amn[1], amn[2], amx[0], amx[1], amx[2]);
            const float v = s + ci * ar;
            h.area[L] = v;
            sah = v;
            my_area = ar;
        }
        const int lastIndex = bcast_half(node, L, L, upper);
        const float sahLast = bcast_half(sah, L, L, upper);
        const float areaLast = bcast_half(my_area, L, L, upper);
        float bmn[3], bmx[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            bmn[k] = box[L * 6 + k];
            bmx[k] = box[L * 6 + 3 + k];
        }
        const int2 spL = sp[L];
        if (hl == r) {
            node = lastIndex;
            sah = sahLast;
            my_area = areaLast;
            sp[r] = spL;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                box[r * 6 + k] = bmn[k];
                box[r * 6 + 3 + k] = bmx[k];
            }
        }
        if (hl == L) {
            node = li;
            h.leaf[L] = ri;
            rec_al = ali;
            rec_ar = ari;
            sp[L] = spm;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                box[L * 6 + k] = amn[k];
                box[L * 6 + 3 + k] = amx[k];
            }
        }
        // updateDistancesMatrix (:152-188)
        if (L > 1) {
            const bool cpy = hl < L && hl != r;
            float cp = 0.0f;
            if (cpy) cp = D[trm(L, hl)];
            if (cpy) D[trm(max(r, hl), min(r, hl))] = cp;
            if (hl < L && hl != c) {
                const int dr = max(c, hl), dc = min(c, hl);
                const float* bi = box + dr * 6;
                const float* bj = box + dc * 6;
                D[trm(dr, dc)] = union_area(bi[0], bi[1], bi[2], bi[3], bi[4], bi[5], bj[0],
                                            bj[1], bj[2], bj[3], bj[4], bj[5]);
            }
        }
    }

    // updateTreelet (:190-212)
    if (h.area[1] < root_sah) {
        if (hl >= 1 && hl < n) {
            const int ni = h.internal[hl - 1];
            const int rr = h.leaf[hl];
            st_coh(&left[ni], node);
            st_coh(&right[ni], rr);
            st_coh(&parent[node], ni);
            st_coh(&parent[rr], ni);
            st_coh(&nsah[ni], h.area[hl]);
            const float4 mn = make_float4(box[hl * 6 + 0], box[hl * 6 + 1], box[hl * 6 + 2], 0.0f);
            const float4 mx = make_float4(box[hl * 6 + 3], box[hl * 6 + 4], box[hl * 6 + 5], 0.0f);
            st_coh(&bbmin[ni], mn);
            st_coh(&bbmax[ni], mx);
            st_coh(&area[ni], box_area4(mn, mx));
            st_coh(&span[ni], sp[hl]);
            st_coh(&rec[ni], make_int4(node, rr, __float_as_int(rec_al), __float_as_int(rec_ar)));
        }
    }
}

// ------------------------------------------------------------------ TRBVH
constexpr int kTrN = 7;
constexpr int kTrSub = 1 << kTrN;
// internal 8 + leaf 8 + area 8 + boxes 48 + leaf spans 16 + subset boxes 768 +
// subset spans 256 + costs 128 + masks 128 + stack 16
constexpr int kTrWords = 8 + 8 + 8 + 48 + 16 + kTrSub * 6 + kTrSub * 2 + kTrSub + kTrSub + 16;

template <bool COH>
__device__ __forceinline__ void trbvh_treelet(float* lds, int hl, int root, float root_sah,
                                              int root_tris, int first_leaf, float ci, float ct,
                                              const int32_t* sched, int rounds, int32_t* parent,
                                              int32_t* left, int32_t* right, float4* bbmin,
                                              float4* bbmax, float* area, float* nsah,
                                              int2* span)
{
    HalfTreelet h{(int32_t*)lds, (int32_t*)lds + 8, lds + 16};
    float* box = lds + 24;
    int2* lsp = (int2*)(box + 48);
    float* sbox = box + 48 + 16;
    int2* ssp = (int2*)(sbox + kTrSub * 6);
    float* cost = sbox + kTrSub * 6 + kTrSub * 2;
    int32_t* mask = (int32_t*)(cost + kTrSub);
    int32_t* stk = mask + kTrSub;
    // formTreelet writes at most index n-1 = 6 into the 8-entry arrays
    form_treelet<COH>(h, hl, root, kTrN, first_leaf, left, right, area);

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
        float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
        float mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        int2 ss = make_int2(0x7fffffff, -1);
        for (int i = 0; i < kTrN; ++i) {
            if (s & (1 << i)) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    mn[k] = fminf(box[i * 6 + k], mn[k]);
                    mx[k] = fmaxf(box[i * 6 + 3 + k], mx[k]);
                }
                ss.x = min(ss.x, lsp[i].x);
                ss.y = max(ss.y, lsp[i].y);
            }
        }
        ssp[s] = ss;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            sbox[s * 6 + k] = mn[k];
            sbox[s * 6 + 3 + k] = mx[k];
        }
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
                int child;
                if (lint) {
                    const int slot = __popc(bl & ltm);
                    child = h.internal[next + slot];
                    const float4 mn = make_float4(sbox[P * 6 + 0], sbox[P * 6 + 1], sbox[P * 6 + 2], 0.0f);
                    const float4 mx = make_float4(sbox[P * 6 + 3], sbox[P * 6 + 4], sbox[P * 6 + 5], 0.0f);
                    st_coh(&bbmin[child], mn);
                    st_coh(&bbmax[child], mx);
                    st_coh(&area[child], box_area4(mn, mx));
                    st_coh(&nsah[child], cost[P]);
                    st_coh(&span[child], ssp[P]);
                    stk[slot] = child;
                    stk[8 + slot] = P;
                } else {
                    child = h.leaf[__ffs(P) - 1];
                }
                st_coh(&left[sroot], child);
                st_coh(&parent[child], sroot);
                if (rint) {
                    const int slot = nl + __popc(br & ltm);
                    child = h.internal[next + slot];
                    const float4 mn = make_float4(sbox[Q * 6 + 0], sbox[Q * 6 + 1], sbox[Q * 6 + 2], 0.0f);
                    const float4 mx = make_float4(sbox[Q * 6 + 3], sbox[Q * 6 + 4], sbox[Q * 6 + 5], 0.0f);
                    st_coh(&bbmin[child], mn);
                    st_coh(&bbmax[child], mx);
                    st_coh(&area[child], box_area4(mn, mx));
                    // the reference stores costs[partition] for the right child too (TRBVH.cu:309)
                    st_coh(&nsah[child], cost[P]);
                    st_coh(&span[child], ssp[Q]);
                    stk[slot] = child;
                    stk[8 + slot] = Q;
                } else {
                    child = h.leaf[__ffs(Q) - 1];
                }
                st_coh(&right[sroot], child);
                st_coh(&parent[child], sroot);
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

template <bool AGG, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_treelet_pass(uint32_t n_leaves, int first_leaf, int n,
                                                      int gamma, float ci,
                                                      float ct, int32_t* parent, int32_t* left,
                                                      int32_t* right, float4* bbmin,
                                                      float4* bbmax, float* area, int32_t* tris,
                                                      float* nsah, uint32_t* counters,
                                                      int2* span, int4* rec,
                                                      const int32_t* sched, int rounds,
                                                      int half_words)
{
    constexpr int kShift = PassShift<BLOCK>::value;
    extern __shared__ float smem[];
    __shared__ uint32_t s_cnt[BLOCK];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int half = lane >> 5;
    const int hl = lane & 31;
    float* lds = smem + ((tid >> 6) * 2 + half) * half_words;

    const uint32_t gid = blockIdx.x * (uint32_t)BLOCK + (uint32_t)tid;
    const int blk = (int)blockIdx.x;
    s_cnt[tid] = 0u;
    __syncthreads();
    int cur = -1;
    if (gid < n_leaves) {
        const int leaf = first_leaf + (int)gid;
        st_coh(&tris[leaf], 1);
        st_coh(&nsah[leaf], ct * area[leaf]);
        cur = parent[leaf];
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
            // meet in global memory (cross-XCD hand-over, see device.hpp).  A
            // node's own links, area, span and parent are unchanged from the
            // start of the launch until the node is processed (only treelets
            // rooted at it or above rewrite them), so they are read plainly and
            // before the arrival: their latency overlaps the arrival's wait for
            // this thread's earlier stores.
            const int2 sp = span[cur];
            const int l = left[cur];
            const int r = right[cur];
            const float ar = area[cur];
            up = parent[cur];
            local = block_local<kShift>(sp, cur, blk);
            const uint32_t old =
                local ? arrive_local(&s_cnt[cur & (BLOCK - 1)])
                      : arrive_crossing(&counters[cur]);
            if (old == 0u) {
                cur = -1;
            } else {
                int tl, tr;
                float sl, sr;
                if (local) {
                    tl = ld<false>(&tris[l]);
                    sl = ld<false>(&nsah[l]);
                    tr = ld<false>(&tris[r]);
                    sr = ld<false>(&nsah[r]);
                } else {
                    tl = ld<true>(&tris[l]);
                    sl = ld<true>(&nsah[l]);
                    tr = ld<true>(&tris[r]);
                    sr = ld<true>(&nsah[r]);
                }
                tc = tl + tr;
                st_coh(&tris[cur], tc);
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
            if (my_r >= 0) {
                if (AGG) {
                    if (wave_local)
                        atrbvh_treelet<false>(lds, hl, root, root_sah, n, first_leaf, ci, parent, left,
                                              right, bbmin, bbmax, area, nsah, span, rec);
                    else
                        atrbvh_treelet<true>(lds, hl, root, root_sah, n, first_leaf, ci, parent, left,
                                             right, bbmin, bbmax, area, nsah, span, rec);
                } else {
                    if (wave_local)
                        trbvh_treelet<false>(lds, hl, root, root_sah, root_tris, first_leaf, ci, ct, sched,
                                             rounds, parent, left, right, bbmin, bbmax, area,
                                             nsah, span);
                    else
                        trbvh_treelet<true>(lds, hl, root, root_sah, root_tris, first_leaf, ci, ct, sched,
                                            rounds, parent, left, right, bbmin, bbmax, area,
                                            nsah, span);
                }
            }
        }
        if (cur >= 0 && !held) cur = up;
    }
}

__global__ void k_zero_u32(uint32_t* p, uint64_t count)
{
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < count; i += 256ull * gridDim.x)
        p[i] = 0u;
}

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

atr_status zero_counters(atr_context* ctx, uint32_t* counters, uint32_t count)
{
    return launch(ctx, "zero_counters", count * 4.0, [&] {
        k_zero_u32<<<std::min<uint32_t>(std::max(div_up(count, 256), 1u), 2048u), 256, 0,
                     ctx->stream>>>(counters, count);
    });
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
                        int32_t* tris, float* nsah, uint32_t* counters, int4* rec,
                        const int32_t* sched, int rounds, int half_words)
{
    const uint32_t N = t->n;
    const size_t lds = (size_t)half_words * (BLOCK / 32) * sizeof(float);
    k_treelet_pass<AGG, BLOCK><<<div_up(N, BLOCK), BLOCK, lds, ctx->stream>>>(
        N, (int)t->first_leaf, n, gamma, ci, ct, t->parent, t->left, t->right, t->bbmin,
        t->bbmax, t->area, tris, nsah, counters, t->span, rec, sched, rounds, half_words);
}

template <bool AGG>
static atr_status launch_pass_block(atr_context* ctx, const char* name, int block, atr_tree* t,
                                    int n, int gamma, float ci, float ct, int32_t* tris,
                                    float* nsah, uint32_t* counters, int4* rec,
                                    const int32_t* sched, int rounds, int half_words)
{
    if (block != 64 && block != 128 && block != 256)
        return fail(ATR_ERR_INVALID_ARGUMENT, "treelet pass block must be 64, 128 or 256");
    return launch(ctx, name, 0.0, [&] {
        if (block == 64)
            launch_pass<AGG, 64>(ctx, t, n, gamma, ci, ct, tris, nsah, counters, rec, sched,
                                 rounds, half_words);
        else if (block == 128)
            launch_pass<AGG, 128>(ctx, t, n, gamma, ci, ct, tris, nsah, counters, rec, sched,
                                  rounds, half_words);
        else
            launch_pass<AGG, 256>(ctx, t, n, gamma, ci, ct, tris, nsah, counters, rec, sched,
                                  rounds, half_words);
    });
}

atr_status run_atrbvh_pass(atr_context* ctx, atr_tree* t, int treelet, int gamma, float ci,
                           float ct, int32_t* tris, float* nsah, uint32_t* counters, int4* rec,
                           int block)
{
    ATR_TRY(zero_counters(ctx, counters, t->first_leaf));
    const int half_words = kAggFixedWords + treelet * (treelet - 1) / 2;
    return launch_pass_block<true>(ctx, "atrbvh_pass", block, t, treelet, gamma, ci, ct, tris,
                                   nsah, counters, rec, nullptr, 0, half_words);
}

atr_status run_trbvh_pass(atr_context* ctx, atr_tree* t, int gamma, float ci, float ct,
                          int32_t* tris, float* nsah, uint32_t* counters,
                          const int32_t* d_schedule, int rounds, int block)
{
    ATR_TRY(zero_counters(ctx, counters, t->first_leaf));
    return launch_pass_block<false>(ctx, "trbvh_pass", block, t, kTrN, gamma, ci, ct, tris, nsah,
                                    counters, nullptr, d_schedule, rounds, kTrWords);
}

}  // namespace atr
