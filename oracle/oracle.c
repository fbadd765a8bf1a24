/*
 * Serial CPU restatement of the BVHRT-Core build path (test infrastructure
 * only -- the checker, never the product).  Compiled with -ffp-contract=off.
 * Each function cites the reference file:line it restates; the header lists
 * the closed-form / brute-force check that pins it.
 */
#include "oracle.h"

/* Test-only: ORC_MUTATION = k plants mistake k (see the #if blocks) so the
 * composition tests can show they catch it; 0 in every normal build. */
#ifndef ORC_MUTATION
#define ORC_MUTATION 0
#endif

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MIN(a, b) fminf((a), (b))
#define ORC_MAX(a, b) fmaxf((a), (b))

static int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }
static int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }
static int popc32(uint32_t x) { return __builtin_popcount(x); }
static int ffs32(uint32_t x) { return x ? __builtin_ctz(x) + 1 : 0; }

static int is_leaf(int index, uint32_t n) { return !((uint32_t)index < n - 1u); }

/* 2*(x*y + x*z + y*z) of a box (Commons.cuh:99-125). */
static float box_area(const float* mn, const float* mx)
{
    float sx = mx[0] - mn[0];
    float sy = mx[1] - mn[1];
    float sz = mx[2] - mn[2];
    return 2.0f * (sx * sy + sx * sz + sy * sz);
}

/* Area of the union of two boxes (Commons.cuh:156-164). */
static float union_area(const float* mn1, const float* mx1, const float* mn2, const float* mx2)
{
    float sx = ORC_MAX(mx1[0], mx2[0]) - ORC_MIN(mn1[0], mn2[0]);
    float sy = ORC_MAX(mx1[1], mx2[1]) - ORC_MIN(mn1[1], mn2[1]);
    float sz = ORC_MAX(mx1[2], mx2[2]) - ORC_MIN(mn1[2], mn2[2]);
    return 2.0f * (sx * sy + sx * sz + sy * sz);
}

/* --------------------------------------------------------------------- */
/* Scene bounds: SceneConverter::ConvertToCudaSceneDeviceMemory
 * (gpu-ray-traversal/src/rt/SceneConverter.cpp:28-64, FindMinAndMax :73-81).
 * Keeps the quirk that the running maximum starts at FLT_MIN (+1.18e-38). */
void orc_scene_bbox(const float* v4, uint32_t n, float bbmin[3], float bbmax[3])
{
    for (int a = 0; a < 3; ++a) {
        bbmin[a] = FLT_MAX;
        bbmax[a] = FLT_MIN;
    }
    for (uint64_t i = 0; i < 3ull * n; ++i) {
        for (int a = 0; a < 3; ++a) {
            float s = v4[i * 4 + a];
            bbmin[a] = (s < bbmin[a]) ? s : bbmin[a]; /* std::min(min, s) */
            bbmax[a] = (bbmax[a] < s) ? s : bbmax[a]; /* std::max(max, s) */
        }
    }
}

/* --------------------------------------------------------------------- */
/* Morton codes: expandBits (Commons.cuh:481-488), calculateMortonCode
 * (:497-510), expandBits64 (:606-618), calculateMortonCode64 (:627-640).
 * The interleave is x*4 + y*2 + z (x most significant). */
uint32_t orc_expand_bits10(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

uint64_t orc_expand_bits21(uint64_t v)
{
    v &= 0x1fffffull;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

uint32_t orc_morton30(float x, float y, float z)
{
    uint32_t qx = (uint32_t)ORC_MIN(ORC_MAX(x * 1024.0f, 0.0f), 1023.0f);
    uint32_t qy = (uint32_t)ORC_MIN(ORC_MAX(y * 1024.0f, 0.0f), 1023.0f);
    uint32_t qz = (uint32_t)ORC_MIN(ORC_MAX(z * 1024.0f, 0.0f), 1023.0f);
    return orc_expand_bits10(qx) * 4u + orc_expand_bits10(qy) * 2u + orc_expand_bits10(qz);
}

uint64_t orc_morton63(float x, float y, float z)
{
    uint64_t qx = (uint64_t)ORC_MIN(ORC_MAX(x * 2097152.0f, 0.0f), 2097151.0f);
    uint64_t qy = (uint64_t)ORC_MIN(ORC_MAX(y * 2097152.0f, 0.0f), 2097151.0f);
    uint64_t qz = (uint64_t)ORC_MIN(ORC_MAX(z * 2097152.0f, 0.0f), 2097151.0f);
    return orc_expand_bits21(qx) * 4ull + orc_expand_bits21(qy) * 2ull + orc_expand_bits21(qz);
}

/* Triangle AABB centroid (LBVH.cu:14-40) normalised by the scene box
 * (Commons.cuh:426-434). */
static void normalized_centroid(const float* v4, uint32_t i, const float* bbmin,
                                const float* bbmax, float* out)
{
    const float* a = v4 + (uint64_t)i * 12;
    const float* b = a + 4;
    const float* c = a + 8;
    for (int k = 0; k < 3; ++k) {
        float mn = ORC_MIN(ORC_MIN(a[k], b[k]), c[k]);
        float mx = ORC_MAX(ORC_MAX(a[k], b[k]), c[k]);
        float cen = (mx + mn) * 0.5f;
        out[k] = (cen - bbmin[k]) / (bbmax[k] - bbmin[k]);
    }
}

/* generateMortonCodesKernel (LBVH.cu:53-74). */
void orc_encode30(const float* v4, uint32_t n, const float bbmin[3], const float bbmax[3],
                  uint32_t* keys, uint32_t* idx)
{
    for (uint32_t i = 0; i < n; ++i) {
        float c[3];
        normalized_centroid(v4, i, bbmin, bbmax, c);
        keys[i] = orc_morton30(c[0], c[1], c[2]);
        if (idx) idx[i] = i;
    }
}

/* generateMortonCodes64Kernel (LBVH.cu:76-97). */
void orc_encode63(const float* v4, uint32_t n, const float bbmin[3], const float bbmax[3],
                  uint64_t* keys, uint32_t* idx)
{
    for (uint32_t i = 0; i < n; ++i) {
        float c[3];
        normalized_centroid(v4, i, bbmin, bbmax, c);
        keys[i] = orc_morton63(c[0], c[1], c[2]);
        if (idx) idx[i] = i;
    }
}

/* --------------------------------------------------------------------- */
/* DeviceSort (CubWrapper.cu:12-62): stable ascending sort of (key, value).
 * Restated as a serial stable LSD counting sort with 8-bit digits. */
void orc_sort_pairs32(uint32_t* keys, uint32_t* vals, uint32_t n)
{
    uint32_t* tk = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* tv = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (int pass = 0; pass < 4; ++pass) {
        uint32_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        int sh = pass * 8;
        for (uint32_t i = 0; i < n; ++i) cnt[((keys[i] >> sh) & 255u) + 1]++;
        for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t p = cnt[(keys[i] >> sh) & 255u]++;
            tk[p] = keys[i];
            tv[p] = vals[i];
        }
        memcpy(keys, tk, sizeof(uint32_t) * n);
        memcpy(vals, tv, sizeof(uint32_t) * n);
    }
    free(tk);
    free(tv);
}

void orc_sort_pairs64(uint64_t* keys, uint32_t* vals, uint32_t n)
{
    uint64_t* tk = (uint64_t*)malloc(sizeof(uint64_t) * (n ? n : 1));
    uint32_t* tv = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (int pass = 0; pass < 8; ++pass) {
        uint32_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        int sh = pass * 8;
        for (uint32_t i = 0; i < n; ++i) cnt[((keys[i] >> sh) & 255u) + 1]++;
        for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t p = cnt[(keys[i] >> sh) & 255u]++;
            tk[p] = keys[i];
            tv[p] = vals[i];
        }
        memcpy(keys, tk, sizeof(uint64_t) * n);
        memcpy(vals, tv, sizeof(uint32_t) * n);
    }
    free(tk);
    free(tv);
}

/* --------------------------------------------------------------------- */
/* Karras radix tree: longestCommonPrefix (LBVH.cu:163-201), buildTreeKernel
 * (LBVH.cu:208-316). Equal keys fall back to the index prefix. */
static int lcp32(const uint32_t* k, uint32_t n, int i, int j, uint32_t ki)
{
    if (j < 0 || j >= (int)n) return 0;
    uint32_t kj = k[j];
    if (ki == kj) return 32 + clz32((uint32_t)(i ^ j));
    return clz32(ki ^ kj);
}

static int lcp64(const uint64_t* k, uint32_t n, int i, int j, uint64_t ki)
{
    if (j < 0 || j >= (int)n) return 0;
    uint64_t kj = k[j];
    if (ki == kj) return 64 + clz64((uint64_t)(int64_t)(i ^ j));
    return clz64(ki ^ kj);
}

static int sgn(int x) { return (0 < x) - (0 > x); }

#define KARRAS_BODY(LCP)                                                                 \
    for (int i = 0; i < (int)n - 1; ++i) {                                               \
        int lcp1 = LCP(keys, n, i, i + 1, keys[i]);                                      \
        int lcp2 = LCP(keys, n, i, i - 1, keys[i]);                                      \
        int dir = sgn(lcp1 - lcp2);                                                      \
        int minLcp = LCP(keys, n, i, i - dir, keys[i]);                                  \
        int lMax = 128;                                                                  \
        while (LCP(keys, n, i, i + lMax * dir, keys[i]) > minLcp) lMax *= 4;             \
        int l = 0;                                                                       \
        int tt = lMax;                                                                   \
        while (tt > 1) {                                                                 \
            tt = tt / 2;                                                                 \
            if (LCP(keys, n, i, i + (l + tt) * dir, keys[i]) > minLcp) l += tt;          \
        }                                                                                \
        int j = i + l * dir;                                                             \
        int nodeLcp = LCP(keys, n, i, j, keys[i]);                                       \
        int s = 0;                                                                       \
        int divisor = 2;                                                                 \
        int maxDivisor = 1 << (32 - clz32((uint32_t)l));                                 \
        while (divisor <= maxDivisor) {                                                  \
            tt = (l + divisor - 1) / divisor;                                            \
            if (LCP(keys, n, i, i + (s + tt) * dir, keys[i]) > nodeLcp) s += tt;         \
            divisor *= 2;                                                                \
        }                                                                                \
        int split = i + s * dir + (dir < 0 ? dir : 0);                                   \
        int li, ri;                                                                      \
        int lo = i < j ? i : j, hi = i < j ? j : i;                                      \
        if (lo == split) {                                                               \
            li = split + (int)n - 1;                                                     \
            t->data[li] = (int32_t)idx[split];                                           \
        } else {                                                                         \
            li = split;                                                                  \
        }                                                                                \
        if (hi == split + 1) {                                                           \
            ri = split + 1 + (int)n - 1;                                                 \
            t->data[ri] = (int32_t)idx[split + 1];                                       \
        } else {                                                                         \
            ri = split + 1;                                                              \
        }                                                                                \
        t->right[i] = ri;                                                                \
        t->left[i] = li;                                                                 \
        t->parent[li] = i;                                                               \
        t->parent[ri] = i;                                                               \
        t->data[i] = -1;                                                                 \
        if (i == 0) {                                                                    \
            t->parent[0] = -1;                                                           \
            t->root = 0;                                                                 \
        }                                                                                \
    }

void orc_karras32(const uint32_t* keys, const uint32_t* idx, uint32_t n, orc_tree* t)
{
    t->n = n;
    KARRAS_BODY(lcp32)
}

void orc_karras64(const uint64_t* keys, const uint32_t* idx, uint32_t n, orc_tree* t)
{
    t->n = n;
    KARRAS_BODY(lcp64)
}

/* --------------------------------------------------------------------- */
/* Iterative post-order walker: calls visit(node) for every node after both of
 * its children.  Children are read when a node is first entered; the
 * bottom-up passes of the reference only ever modify a node's subtree after
 * the node itself has been reached, so this equals its counter-driven order
 * up to the interleaving of disjoint subtrees. */
typedef struct {
    int32_t node;
    int32_t state;
} orc_frame;

typedef void (*orc_visit_fn)(void* ctx, int32_t node);

static int postorder(orc_tree* t, void* ctx, orc_visit_fn visit)
{
    uint32_t n = t->n;
    size_t cap = 64;
    orc_frame* st = (orc_frame*)malloc(sizeof(orc_frame) * cap);
    if (!st) return -1;
    size_t sp = 0;
    st[sp].node = t->root;
    st[sp].state = 0;
    sp++;
    while (sp) {
        orc_frame* f = &st[sp - 1];
        int32_t node = f->node;
        if (is_leaf(node, n) || f->state == 1) {
            sp--;
            visit(ctx, node);
            continue;
        }
        f->state = 1;
        int32_t l = t->left[node], r = t->right[node];
        if (sp + 2 > cap) {
            cap *= 2;
            orc_frame* ns = (orc_frame*)realloc(st, sizeof(orc_frame) * cap);
            if (!ns) {
                free(st);
                return -1;
            }
            st = ns;
        }
        st[sp].node = r;
        st[sp].state = 0;
        sp++;
        st[sp].node = l;
        st[sp].state = 0;
        sp++;
    }
    free(st);
    return 0;
}

/* calculateNodeBoundingBoxesKernel (LBVH.cu:465-570): leaf boxes from the
 * three vertices (calculateLeafBoundingBox :430-446), unions bottom-up
 * (calculateNodeBoundingBox :448-463), areas (Commons.cuh:117-125). */
typedef struct {
    orc_tree* t;
    const float* v4;
} refit_ctx;

static void refit_visit(void* p, int32_t node)
{
    refit_ctx* c = (refit_ctx*)p;
    orc_tree* t = c->t;
    float* mn = t->bbmin + (size_t)node * 4;
    float* mx = t->bbmax + (size_t)node * 4;
    if (is_leaf(node, t->n)) {
        const float* a = c->v4 + (size_t)t->data[node] * 12;
        const float* b = a + 4;
        const float* d = a + 8;
        for (int k = 0; k < 3; ++k) {
            mn[k] = ORC_MIN(ORC_MIN(a[k], b[k]), d[k]);
            mx[k] = ORC_MAX(ORC_MAX(a[k], b[k]), d[k]);
        }
        t->left[node] = -1;
        t->right[node] = -1;
    } else {
        const float* lmn = t->bbmin + (size_t)t->left[node] * 4;
        const float* lmx = t->bbmax + (size_t)t->left[node] * 4;
        const float* rmn = t->bbmin + (size_t)t->right[node] * 4;
        const float* rmx = t->bbmax + (size_t)t->right[node] * 4;
        for (int k = 0; k < 3; ++k) {
            mn[k] = ORC_MIN(lmn[k], rmn[k]);
            mx[k] = ORC_MAX(lmx[k], rmx[k]);
        }
    }
    mn[3] = 0.0f;
    mx[3] = 0.0f;
    t->area[node] = box_area(mn, mx);
}

void orc_refit(const float* v4, orc_tree* t)
{
    refit_ctx c = {t, v4};
    postorder(t, &c, refit_visit);
}

/* LBVHBuilder::BuildTreeNoTriangleSplitting (LBVHBuilder.cpp:36-130). */
int orc_lbvh(const float* v4, uint32_t n, const float bbmin[3], const float bbmax[3],
             int morton_bits, orc_tree* t)
{
    if (n < 2) return -1;
    uint32_t* idx = (uint32_t*)malloc(sizeof(uint32_t) * n);
    if (!idx) return -1;
    t->n = n;
    if (morton_bits == 30) {
        uint32_t* k = (uint32_t*)malloc(sizeof(uint32_t) * n);
        if (!k) {
            free(idx);
            return -1;
        }
        orc_encode30(v4, n, bbmin, bbmax, k, idx);
        orc_sort_pairs32(k, idx, n);
        orc_karras32(k, idx, n, t);
        free(k);
    } else if (morton_bits == 63) {
        uint64_t* k = (uint64_t*)malloc(sizeof(uint64_t) * n);
        if (!k) {
            free(idx);
            return -1;
        }
        orc_encode63(v4, n, bbmin, bbmax, k, idx);
        orc_sort_pairs64(k, idx, n);
        orc_karras64(k, idx, n, t);
        free(k);
    } else {
        free(idx);
        return -1;
    }
    free(idx);
    orc_refit(v4, t);
    return 0;
}

/* The triangle-splitting LBVH over primitive boxes (LBVH.cu): keys of the
 * box centroids (generateMortonCodesTSKernel / generateMortonCodes64TSKernel
 * :99-161, getBoundingBoxCentroid :42-51) with identity routing indices,
 * the stable sort, buildTreeTSKernel (:320-428: leaf data index =
 * tri[routing]) and calculateNodeBoundingBoxesTSKernel (:572-669: leaf box =
 * the routed box; w set to 0 here).  bmin4/bmax4 are float4[m]. */
typedef struct {
    orc_tree* t;
    const float* bmin4;
    const float* bmax4;
} box_refit_ctx;

static void box_refit_visit(void* p, int32_t node)
{
    box_refit_ctx* c = (box_refit_ctx*)p;
    orc_tree* t = c->t;
    float* mn = t->bbmin + (size_t)node * 4;
    float* mx = t->bbmax + (size_t)node * 4;
    if (is_leaf(node, t->n)) {
        /* data holds the routing index until the build remaps it */
        const float* a = c->bmin4 + (size_t)t->data[node] * 4;
        const float* b = c->bmax4 + (size_t)t->data[node] * 4;
        for (int k = 0; k < 3; ++k) {
            mn[k] = a[k];
            mx[k] = b[k];
        }
        t->left[node] = -1;
        t->right[node] = -1;
    } else {
        const float* lmn = t->bbmin + (size_t)t->left[node] * 4;
        const float* lmx = t->bbmax + (size_t)t->left[node] * 4;
        const float* rmn = t->bbmin + (size_t)t->right[node] * 4;
        const float* rmx = t->bbmax + (size_t)t->right[node] * 4;
        for (int k = 0; k < 3; ++k) {
            mn[k] = ORC_MIN(lmn[k], rmn[k]);
            mx[k] = ORC_MAX(lmx[k], rmx[k]);
        }
    }
    mn[3] = 0.0f;
    mx[3] = 0.0f;
    t->area[node] = box_area(mn, mx);
}

int orc_lbvh_boxes(const float* bmin4, const float* bmax4, const int32_t* tri, uint32_t m,
                   const float bbmin[3], const float bbmax[3], int morton_bits, orc_tree* t)
{
    if (m < 2 || (morton_bits != 30 && morton_bits != 63)) return -1;
    uint32_t* idx = (uint32_t*)malloc(sizeof(uint32_t) * m);
    void* k = malloc((morton_bits == 63 ? 8 : 4) * (size_t)m);
    if (!idx || !k) {
        free(idx);
        free(k);
        return -1;
    }
    t->n = m;
    for (uint32_t i = 0; i < m; ++i) {
        float c[3];
        for (int a = 0; a < 3; ++a) {
            const float cen = (bmin4[(size_t)i * 4 + a] + bmax4[(size_t)i * 4 + a]) / 2.0f;
            c[a] = (cen - bbmin[a]) / (bbmax[a] - bbmin[a]);
        }
        if (morton_bits == 30) ((uint32_t*)k)[i] = orc_morton30(c[0], c[1], c[2]);
        else ((uint64_t*)k)[i] = orc_morton63(c[0], c[1], c[2]);
        idx[i] = i;
    }
    if (morton_bits == 30) {
        orc_sort_pairs32((uint32_t*)k, idx, m);
        orc_karras32((uint32_t*)k, idx, m, t);
    } else {
        orc_sort_pairs64((uint64_t*)k, idx, m);
        orc_karras64((uint64_t*)k, idx, m, t);
    }
    free(k);
    free(idx);
    box_refit_ctx c = {t, bmin4, bmax4};
    if (postorder(t, &c, box_refit_visit) != 0) return -1;
    for (uint32_t i = 0; i < m; ++i) {
        const size_t leaf = (size_t)m - 1 + i;
        t->data[leaf] = tri[t->data[leaf]];
    }
    return 0;
}

/* SoABVHTree::SAH (SoABVHTree.cpp:50-74): serial float sums, root first,
 * then the other internal nodes in index order, then leaves. */
float orc_sah(const orc_tree* t, float ci, float ct)
{
    int root = t->root;
    float total = t->area[root];
    float sumInternal = total;
    float sumLeaves = 0.0f;
    for (uint32_t i = 0; i < t->n - 1; ++i)
        if ((int)i != root) sumInternal += t->area[i];
    for (uint32_t i = t->n - 1; i < 2 * t->n - 1; ++i) sumLeaves += t->area[i];
    return (ci * sumInternal + ct * sumLeaves) / total;
}

/* --------------------------------------------------------------------- */
/* LOWER_TRM_ROW / LOWER_TRM_COL (TriangularMatrix.cuh:40-52). */
int orc_lower_trm_row(int index) { return (int)((-1.0f + sqrtf(8.0f * (float)index + 1.0f)) / 2.0f) + 1; }

int orc_lower_trm_col(int index)
{
    int y = orc_lower_trm_row(index);
    return index - (y * (y - 1)) / 2;
}

#define TRM_INDEX(row, col) ((((row) - 1) * (row)) / 2 + (col))

/* --------------------------------------------------------------------- */
/* TRBVHScheduler::GenerateSchedule / ProcessSubset
 * (TRBVHScheduler.cpp:29-126).  Dependencies are the supersets a subset was
 * partitioned from; subsets of each size are taken in ascending value order
 * (std::set order), placed in the earliest round after all dependencies with
 * at most `warp` entries per round, padded with 0 and the rounds reversed. */
typedef struct {
    int n;
    uint64_t* deps; /* (1<<n) x 2 words: bitmap of supersets (n <= 7) */
    uint8_t* expanded;
} sched_ctx;

static void sched_process(sched_ctx* c, int subset, int superset)
{
    int size = popc32((uint32_t)subset);
    if (size == 1) return;
    if (size <= c->n - 2) c->deps[subset * 2 + (superset >> 6)] |= 1ull << (superset & 63);
    if (size > 2 && !c->expanded[subset]) {
        c->expanded[subset] = 1;
        int delta = (subset - 1) & subset;
        int partition = (-delta) & subset;
        while (partition != 0) {
            sched_process(c, partition, subset);
            sched_process(c, partition ^ subset, subset);
            partition = (partition - delta) & subset;
        }
    }
}

int orc_trbvh_schedule(int treelet, int warp, int32_t* out, int cap)
{
    if (treelet < 3 || treelet > 7 || warp <= 0) return -1;
    int nsub = 1 << treelet;
    sched_ctx c;
    c.n = treelet;
    c.deps = (uint64_t*)calloc((size_t)nsub * 2, sizeof(uint64_t));
    c.expanded = (uint8_t*)calloc((size_t)nsub, 1);
    int* round_of = (int*)malloc(sizeof(int) * nsub);
    int max_rounds = nsub + 1;
    int* fill = (int*)calloc((size_t)max_rounds, sizeof(int));
    int32_t* rounds = (int32_t*)calloc((size_t)max_rounds * warp, sizeof(int32_t));
    if (!c.deps || !c.expanded || !round_of || !fill || !rounds) {
        free(c.deps);
        free(c.expanded);
        free(round_of);
        free(fill);
        free(rounds);
        return -1;
    }
    for (int s = 0; s < nsub; ++s) round_of[s] = -1;
    sched_process(&c, nsub - 1, 0);
    int nrounds = 0;
    for (int size = treelet - 2; size >= 2; --size) {
        for (int s = 1; s < nsub; ++s) {
            if (popc32((uint32_t)s) != size) continue;
            unsigned minimum = 0;
            for (int d = 0; d < nsub; ++d) {
                if (!(c.deps[s * 2 + (d >> 6)] >> (d & 63) & 1ull)) continue;
                unsigned dr = (unsigned)round_of[d] + 1u;
                if (dr > minimum) minimum = dr;
            }
            unsigned r = minimum;
            while (fill[r] == warp) ++r;
            rounds[(size_t)r * warp + fill[r]++] = s;
            round_of[s] = (int)r;
            if ((int)r + 1 > nrounds) nrounds = (int)r + 1;
        }
    }
    int ok = nrounds * warp <= cap;
    if (ok) {
        for (int r = 0; r < nrounds; ++r)
            memcpy(out + (size_t)r * warp, rounds + (size_t)(nrounds - 1 - r) * warp,
                   sizeof(int32_t) * warp);
    }
    free(c.deps);
    free(c.expanded);
    free(round_of);
    free(fill);
    free(rounds);
    return ok ? nrounds : -1;
}

/* AgglomerativeScheduler::GenerateScheduleLower (AgglomerativeScheduler.cpp:46-85):
 * all pairs (i > j) packed two per 32-bit word as (i << 8 | j) halves. */
int orc_atrbvh_schedule(int treelet, int warp, int32_t* out, int cap)
{
    if (treelet < 2 || warp <= 0) return -1;
    int elements = (treelet - 1) * treelet / 2;
    int per = 2 * warp;
    int iters = (elements + per - 1) / per;
    int size = iters * warp;
    if (size > cap) return -1;
    for (int i = 0; i < size; ++i) out[i] = 0;
    int count = 0;
    float multiplier = 0.0f;
    for (int i = 0; i < treelet; ++i) {
        for (int j = 0; j < i; ++j) {
            int index = count + (int)multiplier * warp;
            out[index] = (int32_t)(((uint32_t)out[index] << 16) | (uint32_t)(i << 8) | (uint32_t)j);
            ++count;
            if (count == warp) {
                multiplier += 0.5f;
                count = 0;
            }
        }
    }
    if ((int)(multiplier + 0.5f) > (int)multiplier) {
        for (int i = count; i < warp; ++i) {
            int index = i + (int)multiplier * warp;
            out[index] = (int32_t)((uint32_t)out[index] << 16);
        }
    }
    return size;
}

/* --------------------------------------------------------------------- */
/* findLeafToExpand (Treelet.cuh:23-39) emulated over 32 lanes: a halving
 * tree of width m with __shfl_down semantics (reads above lane 31 return the
 * lane's own value), strict '>' so ties keep the lower lane. */
static int emulate_find_leaf_to_expand(int m, const float* area_in)
{
    float a[32], na[32];
    int ix[32], ni[32];
    for (int l = 0; l < 32; ++l) {
        a[l] = l < m ? area_in[l] : -INFINITY;
        ix[l] = l < m ? l : -1;
    }
    int count = m;
    int shift = count / 2;
    while (count > 1) {
        for (int l = 0; l < 32; ++l) {
            int src = l + shift;
            float oa = src < 32 ? a[src] : a[l];
            int oi = src < 32 ? ix[src] : ix[l];
            if (oa > a[l]) {
                na[l] = oa;
                ni[l] = oi;
            } else {
                na[l] = a[l];
                ni[l] = ix[l];
            }
        }
        memcpy(a, na, sizeof(a));
        memcpy(ix, ni, sizeof(ix));
        count = (count + 1) / 2;
        shift = count / 2;
    }
    return ix[0];
}

/* formTreelet (Treelet.cuh:57-114). */
static int form_treelet(orc_tree* t, int root, int n, int32_t* internal, int32_t* leaves,
                        float* areas)
{
    uint32_t N = t->n;
    internal[0] = root;
    leaves[0] = t->left[root];
    leaves[1] = t->right[root];
    areas[0] = t->area[leaves[0]];
    areas[1] = t->area[leaves[1]];
    for (int iter = 0; iter < n - 2; ++iter) {
        int m = 2 + iter;
        float cand[32];
        for (int l = 0; l < m; ++l) cand[l] = is_leaf(leaves[l], N) ? -FLT_MAX : areas[l];
#if ORC_MUTATION == 2 /* planted: the smallest internal candidate is expanded */
        for (int l = 0; l < m; ++l)
            if (!is_leaf(leaves[l], N)) cand[l] = FLT_MAX - areas[l];
#endif
        int chosen = emulate_find_leaf_to_expand(m, cand);
        if (chosen < 0 || chosen >= m || is_leaf(leaves[chosen], N)) return -1;
        int replace = leaves[chosen];
        int l = t->left[replace];
        int r = t->right[replace];
#if ORC_MUTATION == 1 /* planted: the expanded node's children change places */
        { int tmp = l; l = r; r = tmp; }
#endif
        internal[iter + 1] = replace;
        leaves[chosen] = l;
        leaves[iter + 2] = r;
        areas[chosen] = t->area[l];
        areas[iter + 2] = t->area[r];
    }
    return 0;
}

/* --------------------------------------------------------------------- */
/* TRBVH single-treelet DP (TRBVH.cu:100-220) on precomputed inputs.
 * costs_out/masks_out hold 1<<n entries.  Returns the full-set cost. */
static float trbvh_pick_cost(float c, int p, float* best, int* bestp)
{
    if (c < *best) {
        *best = c;
        *bestp = p;
    }
    return c;
}

float orc_trbvh_dp(int n, const float* leaf_cost, const float* subset_area, int treelet_tris,
                   float ci, float ct, float* costs, int32_t* masks)
{
    int nsub = 1 << n;
    int full = nsub - 1;
    float T = (float)treelet_tris;
#if ORC_MUTATION == 3 /* planted: each subset's collapse term uses its own size */
#define ORC_T_OF(s) ((float)popc32((uint32_t)(s)))
#else
#define ORC_T_OF(s) T
#endif
    for (int s = 0; s < nsub; ++s) {
        costs[s] = subset_area[s];
        masks[s] = 0;
    }
    for (int i = 0; i < n; ++i) costs[1 << i] = leaf_cost[i];

    /* processSchedule (TRBVH.cu:100-134): sizes 2..n-2, sequential partition
     * scan in delta order, first strict minimum kept. */
    for (int size = 2; size <= n - 2; ++size) {
        for (int s = 1; s < nsub; ++s) {
            if (popc32((uint32_t)s) != size) continue;
            float best = FLT_MAX;
            int bestp = 0;
            int delta = (s - 1) & s;
            int p = (-delta) & s;
            while (p != 0) {
#if ORC_MUTATION == 6 /* planted: the first partition of each subset is skipped */
                if (p == ((-delta) & s) && ((p - delta) & s) != 0) {
                    p = (p - delta) & s;
                    continue;
                }
#endif
                trbvh_pick_cost(costs[p] + costs[p ^ s], p, &best, &bestp);
                p = (p - delta) & s;
            }
            float a = costs[s];
            costs[s] = ORC_MIN(ci * a + best, ct * a * ORC_T_OF(s));
            masks[s] = (int8_t)bestp;
        }
    }

    /* processSubsets, size n-1 (TRBVH.cu:139-187): four lanes per subset,
     * lane r scans partition bases r+1, r+5, ...; 4-lane halving reduction. */
    for (int k = 0; k < n; ++k) {
        int s = full & ~(1 << k);
        int leftMask = -(1 << (k - 1 > 0 ? k - 1 : 0));
        int parts = (1 << (n - 2)) - 1;
        float lc[4];
        int lp[4];
        for (int r = 0; r < 4; ++r) {
            float best = FLT_MAX;
            int bestp = 0;
            int base = r + 1;
            int p = (((base & leftMask) << 1) | (base & ~leftMask)) << 1;
            for (int j = r; j < parts; j += 4) {
                trbvh_pick_cost(costs[p] + costs[p ^ s], p, &best, &bestp);
                base += 4;
                p = (((base & leftMask) << 1) | (base & ~leftMask)) << 1;
            }
            lc[r] = best;
            lp[r] = bestp;
        }
        /* reduceOptimal(..., 4) (TRBVH.cu:23-35): shift 2 then shift 1. */
        if (lc[2] < lc[0]) {
            lc[0] = lc[2];
            lp[0] = lp[2];
        }
        if (lc[3] < lc[1]) {
            lc[1] = lc[3];
            lp[1] = lp[3];
        }
        if (lc[1] < lc[0]) {
            lc[0] = lc[1];
            lp[0] = lp[1];
        }
        float a = costs[s];
        costs[s] = ORC_MIN(ci * a + lc[0], ct * a * ORC_T_OF(s));
        masks[s] = (int8_t)lp[0];
    }

    /* size n (TRBVH.cu:192-219): 32 lanes, lane l scans 2(l+1) + 64j,
     * then reduceOptimal(..., 32). */
    {
        float lc[32];
        int lp[32];
        int parts = (1 << (n - 1)) - 1;
        for (int l = 0; l < 32; ++l) {
            float best = FLT_MAX;
            int bestp = 0;
            int p = (l + 1) * 2;
            for (int j = l; j < parts; j += 32) {
                trbvh_pick_cost(costs[p] + costs[p ^ full], p, &best, &bestp);
                p += 64;
            }
            lc[l] = best;
            lp[l] = bestp;
        }
        for (int sh = 16; sh > 0; sh >>= 1) {
            float nc[32];
            int np[32];
            for (int l = 0; l < 32; ++l) {
                int src = l + sh;
                float oc = src < 32 ? lc[src] : lc[l];
                int op = src < 32 ? lp[src] : lp[l];
                if (oc < lc[l]) {
                    nc[l] = oc;
                    np[l] = op;
                } else {
                    nc[l] = lc[l];
                    np[l] = lp[l];
                }
            }
            memcpy(lc, nc, sizeof(lc));
            memcpy(lp, np, sizeof(lp));
        }
        float a = costs[full];
        costs[full] = ORC_MIN(ci * a + lc[0], ct * a * ORC_T_OF(full));
        masks[full] = (int8_t)lp[0];
    }
#undef ORC_T_OF
    return costs[full];
}

/* One TRBVH treelet: formTreelet, calculateSubsetSurfaceAreas (TRBVH.cu:37-98),
 * leaf costs (:438-443), DP, updateTreelet (:222-327). */
static int trbvh_treelet(orc_tree* t, int root, int n, const int32_t* tris, float* nsah,
                         float ci, float ct)
{
    int32_t internal[32], leaves[32];
    float areas[32];
    if (form_treelet(t, root, n, internal, leaves, areas)) return -1;
    int nsub = 1 << n;
    float smn[3 * 128], smx[3 * 128], sarea[128], costs[128];
    int32_t masks[128];
    for (int s = 0; s < nsub; ++s) {
        float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
        float mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (int i = 0; i < n; ++i) {
            if (!(s & (1 << i))) continue;
            const float* lmn = t->bbmin + (size_t)leaves[i] * 4;
            const float* lmx = t->bbmax + (size_t)leaves[i] * 4;
            for (int k = 0; k < 3; ++k) {
                mn[k] = ORC_MIN(lmn[k], mn[k]);
                mx[k] = ORC_MAX(lmx[k], mx[k]);
            }
        }
        memcpy(smn + 3 * s, mn, sizeof(mn));
        memcpy(smx + 3 * s, mx, sizeof(mx));
        sarea[s] = box_area(mn, mx);
    }
    float leaf_cost[32];
    for (int i = 0; i < n; ++i) leaf_cost[i] = nsah[leaves[i]];
    orc_trbvh_dp(n, leaf_cost, sarea, tris[internal[0]], ci, ct, costs, masks);

    int full = nsub - 1;
    if (costs[full] < nsah[internal[0]]) {
        int32_t stackNode[32], stackMask[32], nextNode[32], nextMask[32];
        int sp = 1, cur = 1;
        stackNode[0] = internal[0];
        stackMask[0] = full;
        while (sp > 0) {
            int nsp = 0;
            /* left children of every stack entry (lane order), then right. */
            for (int side = 0; side < 2; ++side) {
                for (int e = 0; e < sp; ++e) {
                    int subset = stackMask[e];
                    int p = (int)(int8_t)masks[subset];
                    int part = side == 0 ? p : ((p ^ subset) & 0xFF);
                    int sroot = stackNode[e];
                    int child;
                    if (popc32((uint32_t)part) > 1) {
                        child = internal[cur++];
                        t->parent[child] = sroot;
                        float* mn = t->bbmin + (size_t)child * 4;
                        float* mx = t->bbmax + (size_t)child * 4;
                        memcpy(mn, smn + 3 * part, sizeof(float) * 3);
                        memcpy(mx, smx + 3 * part, sizeof(float) * 3);
                        mn[3] = 0.0f;
                        mx[3] = 0.0f;
                        t->area[child] = box_area(mn, mx);
                        /* both children receive costs[partition] (TRBVH.cu:275,309) */
                        nsah[child] = costs[p];
#if ORC_MUTATION == 7 /* planted: the right child gets its own cost (quirk removed) */
                        if (side == 1) nsah[child] = costs[part];
#endif
                        nextNode[nsp] = child;
                        nextMask[nsp] = part;
                        nsp++;
                    } else {
                        child = leaves[ffs32((uint32_t)part) - 1];
                        t->parent[child] = sroot;
                    }
                    if (side == 0)
                        t->left[sroot] = child;
                    else
                        t->right[sroot] = child;
                }
            }
            memcpy(stackNode, nextNode, sizeof(int32_t) * nsp);
            memcpy(stackMask, nextMask, sizeof(int32_t) * nsp);
            sp = nsp;
        }
#if ORC_MUTATION == 8 /* planted: the rewritten root's SAH is updated (quirk removed) */
        nsah[internal[0]] = costs[full];
#endif
    }
    return 0;
}

/* --------------------------------------------------------------------- */
/* findMinimumDistance (AgglomerativeTreelet.cu:83-99) with findMinimum
 * (:65-81): lane l scans entries l, l+32, ... keeping the first strict
 * minimum, then a 32-lane halving reduction with strict '<'. */
static int emulate_find_minimum(const float* D, int lastRow)
{
    float md[32], nd[32];
    int mi[32], ni[32];
    int size = lastRow * (1 + lastRow) / 2;
    for (int l = 0; l < 32; ++l) {
        md[l] = FLT_MAX;
        mi[l] = 0;
        for (int j = l; j < size; j += 32) {
            if (D[j] < md[l]) {
                md[l] = D[j];
                mi[l] = j;
            }
        }
    }
    for (int sh = 16; sh > 0; sh >>= 1) {
        for (int l = 0; l < 32; ++l) {
            int src = l + sh;
            float o = src < 32 ? md[src] : md[l];
            int oi = src < 32 ? mi[src] : mi[l];
            if (o < md[l]) {
                nd[l] = o;
                ni[l] = oi;
            } else {
                nd[l] = md[l];
                ni[l] = mi[l];
            }
        }
        memcpy(md, nd, sizeof(md));
        memcpy(mi, ni, sizeof(mi));
    }
    return mi[0];
}

/* Test entry points for the two emulated reductions above. */
int orc_find_minimum(const float* D, int lastRow) { return emulate_find_minimum(D, lastRow); }

int orc_find_leaf_to_expand(int m, const float* area)
{
    if (m < 1 || m > 32) return -1;
    return emulate_find_leaf_to_expand(m, area);
}

/* Agglomerative clustering of one treelet (AgglomerativeTreelet.cu:492-532 with
 * calculateDistancesMatrix :18-63, updateState :101-150 and
 * updateDistancesMatrix :152-188).  Lane state: node/sah/box per slot.
 * node_io: in = treelet leaves, out = per-slot left child for slots >= 1;
 * right_out: per-slot right child; areas_out[k] = SAH of the node built at
 * step lastRow = k; boxes_io: per-slot boxes (in leaves, out merged). */
typedef struct {
    int32_t node[32];
    int32_t right[32];
    float sah[32];
    float mn[32][3];
    float mx[32][3];
    float areas[32];
} agg_state;

static void agglomerate(agg_state* S, int n, const int32_t* internal, float ci, int32_t* merges)
{
    float D[496];
    for (int i = 1; i < n; ++i)
        for (int j = 0; j < i; ++j) D[TRM_INDEX(i, j)] = union_area(S->mn[i], S->mx[i], S->mn[j], S->mx[j]);

    for (int L = n - 1; L > 0; --L) {
        int minIndex = emulate_find_minimum(D, L);
        int r = orc_lower_trm_row(minIndex);
        int c = orc_lower_trm_col(minIndex);
        if (merges) {
            merges[(n - 1 - L) * 2] = r;
            merges[(n - 1 - L) * 2 + 1] = c;
        }
        /* updateState */
        float sah = S->sah[r] + S->sah[c];
        int32_t li = S->node[r];
        int32_t ri = S->node[c];
        float amn[3], amx[3];
        for (int k = 0; k < 3; ++k) {
            amn[k] = ORC_MIN(S->mn[c][k], S->mn[r][k]);
            amx[k] = ORC_MAX(S->mx[c][k], S->mx[r][k]);
        }
        S->node[c] = internal ? internal[L - 1] : -(L + 1);
        memcpy(S->mn[c], amn, sizeof(amn));
        memcpy(S->mx[c], amx, sizeof(amx));
        float area = box_area(amn, amx);
        S->areas[L] = sah + ci * area;
#if ORC_MUTATION == 9 /* planted: the merged node's SAH drops its own area term */
        S->areas[L] = sah;
#endif
        S->sah[c] = S->areas[L];

        int32_t lastIndex = S->node[L];
        float sahLast = S->sah[L];
        float bmn[3], bmx[3];
        memcpy(bmn, S->mn[L], sizeof(bmn));
        memcpy(bmx, S->mx[L], sizeof(bmx));
        S->node[r] = lastIndex;
        S->sah[r] = sahLast;
        memcpy(S->mn[r], bmn, sizeof(bmn));
        memcpy(S->mx[r], bmx, sizeof(bmx));

        S->node[L] = li;
        S->right[L] = ri;
        memcpy(S->mn[L], amn, sizeof(amn));
        memcpy(S->mx[L], amx, sizeof(amx));

        if (L > 1) {
            float copy[32];
            for (int tt = 0; tt < L; ++tt)
                if (tt != r) copy[tt] = D[TRM_INDEX(L, tt)];
            for (int tt = 0; tt < L; ++tt) {
                if (tt == r) continue;
                int dr = r > tt ? r : tt, dc = r < tt ? r : tt;
                D[TRM_INDEX(dr, dc)] = copy[tt];
            }
            for (int tt = 0; tt < L; ++tt) {
                if (tt == c) continue;
                int dr = c > tt ? c : tt, dc = c < tt ? c : tt;
#if ORC_MUTATION == 5 /* planted: distances to the merged cluster are not refreshed */
                continue;
#endif
                D[TRM_INDEX(dr, dc)] = union_area(S->mn[dr], S->mx[dr], S->mn[dc], S->mx[dc]);
            }
        }
    }
}

int orc_atrbvh_cluster(int n, const float* boxes, const float* leaf_sah, float ci,
                       int32_t* merges, float* areas_out)
{
    if (n < 2 || n > 32) return -1;
    agg_state S;
    for (int k = 0; k < n; ++k) {
        S.node[k] = k;
        S.right[k] = -1;
        S.sah[k] = leaf_sah[k];
        for (int a = 0; a < 3; ++a) {
            S.mn[k][a] = boxes[k * 6 + a];
            S.mx[k][a] = boxes[k * 6 + 3 + a];
        }
        S.areas[k] = 0.0f;
    }
    agglomerate(&S, n, NULL, ci, merges);
    if (areas_out)
        for (int k = 0; k < n; ++k) areas_out[k] = S.areas[k];
    return 0;
}

/* One ATRBVH treelet (agglomerativeSmallTreeletKernel body,
 * AgglomerativeTreelet.cu:473-546; updateTreelet :190-212). */
static int atrbvh_treelet(orc_tree* t, int root, int n, float* nsah, float ci)
{
    int32_t internal[32], leaves[32];
    float areas[32];
    if (form_treelet(t, root, n, internal, leaves, areas)) return -1;
    agg_state S;
    for (int k = 0; k < n; ++k) {
        S.node[k] = leaves[k];
        S.right[k] = -1;
        S.sah[k] = nsah[leaves[k]];
        for (int a = 0; a < 3; ++a) {
            S.mn[k][a] = t->bbmin[(size_t)leaves[k] * 4 + a];
            S.mx[k][a] = t->bbmax[(size_t)leaves[k] * 4 + a];
        }
        S.areas[k] = areas[k];
    }
    agglomerate(&S, n, internal, ci, NULL);
    if (S.areas[1] < nsah[internal[0]]) {
        for (int k = 1; k < n; ++k) {
            int node = internal[k - 1];
            t->left[node] = S.node[k];
            t->right[node] = S.right[k];
            t->parent[S.node[k]] = node;
            t->parent[S.right[k]] = node;
            nsah[node] = S.areas[k];
            float* mn = t->bbmin + (size_t)node * 4;
            float* mx = t->bbmax + (size_t)node * 4;
            memcpy(mn, S.mn[k], sizeof(float) * 3);
            memcpy(mx, S.mx[k], sizeof(float) * 3);
            mn[3] = 0.0f;
            mx[3] = 0.0f;
            t->area[node] = box_area(mn, mx);
        }
    }
    return 0;
}

/* --------------------------------------------------------------------- */
/* Bottom-up treelet pass shared by both optimisers
 * (treeletReestructureKernel TRBVH.cu:329-479,
 *  agglomerativeSmallTreeletKernel AgglomerativeTreelet.cu:392-554). */
typedef struct {
    orc_tree* t;
    int kind; /* 0 = TRBVH, 1 = ATRBVH */
    int n;
    int gamma;
    float ci, ct;
    int32_t* tris;
    float* nsah;
    int err;
    int treelets; /* treelet roots processed in this pass */
} pass_ctx;

static void pass_visit(void* p, int32_t node)
{
    pass_ctx* c = (pass_ctx*)p;
    orc_tree* t = c->t;
    if (is_leaf(node, t->n)) {
        c->tris[node] = 1;
        c->nsah[node] = c->ct * t->area[node];
        return;
    }
    int l = t->left[node], r = t->right[node];
    int count = c->tris[l] + c->tris[r];
    c->tris[node] = count;
    c->nsah[node] = c->ci * t->area[node] + c->nsah[l] + c->nsah[r];
    if (count >= c->gamma) {
        c->treelets++;
        int e = c->kind == 0 ? trbvh_treelet(t, node, c->n, c->tris, c->nsah, c->ci, c->ct)
                             : atrbvh_treelet(t, node, c->n, c->nsah, c->ci);
        if (e) c->err = e;
    }
}

/* One bottom-up pass with treelet threshold gamma: every node whose subtree
 * holds at least gamma triangles when the pass reaches it roots a treelet.
 * Returns the number of treelets, or -1. */
static int treelet_pass(orc_tree* t, int kind, int n, int gamma, float ci, float ct)
{
    size_t nodes = 2 * (size_t)t->n - 1;
    pass_ctx c;
    c.t = t;
    c.kind = kind;
    c.n = n;
    c.gamma = gamma;
    c.ci = ci;
    c.ct = ct;
    c.err = 0;
    c.treelets = 0;
    c.tris = (int32_t*)malloc(sizeof(int32_t) * nodes);
    c.nsah = (float*)malloc(sizeof(float) * nodes);
    if (!c.tris || !c.nsah) {
        free(c.tris);
        free(c.nsah);
        return -1;
    }
    if (postorder(t, &c, pass_visit)) c.err = -1;
    free(c.tris);
    free(c.nsah);
    return c.err ? -1 : c.treelets;
}

int orc_treelet_pass(orc_tree* t, int kind, int treelet, int gamma, float ci, float ct)
{
    if (kind == 0 && treelet != 7) return -1;
    if (kind == 1 && (treelet < 3 || treelet > 32)) return -1;
    if (t->n < 2 || gamma < 1) return -1;
    return treelet_pass(t, kind, treelet, gamma, ci, ct);
}

/* The optimisers' schedule: iteration it uses gamma = n * 2^it
 * (TRBVHOptimizer.cpp:91,105; AgglomerativeTreeletOptimizer.cpp:79-96). */
static int treelet_optimize(orc_tree* t, int kind, int n, int iterations, float ci, float ct)
{
    int gamma = n;
    for (int it = 0; it < iterations; ++it) {
        if (treelet_pass(t, kind, n, gamma, ci, ct) < 0) return -1;
#if ORC_MUTATION == 4 /* planted: gamma grows linearly */
        gamma += n;
#else
        gamma *= 2;
#endif
    }
    return 0;
}

/* TRBVHOptimizer::Optimize (TRBVHOptimizer.cpp:21-128); treelet size 7 only
 * (subset masks are 8-bit and 4 lanes x n subsets must fit in 32 lanes). */
int orc_trbvh_optimize(orc_tree* t, int treelet, int iterations, float ci, float ct)
{
    if (treelet != 7 || iterations < 0 || t->n < 2) return -1;
    return treelet_optimize(t, 0, treelet, iterations, ci, ct);
}

/* AgglomerativeTreeletOptimizer::Optimize (AgglomerativeTreeletOptimizer.cpp:24-110). */
int orc_atrbvh_optimize(orc_tree* t, int treelet, int iterations, float ci, float ct)
{
    if (treelet < 3 || treelet > 32 || iterations < 0 || t->n < 2) return -1;
    return treelet_optimize(t, 1, treelet, iterations, ci, ct);
}

/* --------------------------------------------------------------------- */
/* BVHTreeCollapser::Collapse (BVHTreeCollapser.cpp:21-93) with copyTreeKernel,
 * calculateCostKernel and collapseTreeKernel (TreeCollapser.cu:13-237).
 * Offsets of emitted leaf lists: emitting parents in ascending node index,
 * left child before right (the reference takes them from an atomicAdd in
 * arbitrary order). */
typedef struct {
    const orc_tree* t;
    orc_collapsed* c;
    float* cost;
    uint8_t* flag;
    uint8_t* deliver;
    float ci, ct;
} cost_ctx;

static void cost_visit(void* p, int32_t node)
{
    cost_ctx* x = (cost_ctx*)p;
    const orc_tree* t = x->t;
    const float* mn = t->bbmin + (size_t)node * 4;
    const float* mx = t->bbmax + (size_t)node * 4;
    if (is_leaf(node, t->n)) {
        x->c->count[node] = 1;
        x->cost[node] = x->ct * box_area(mn, mx);
        x->flag[node] = 1;
        x->deliver[node] = 1;
        return;
    }
    int l = t->left[node], r = t->right[node];
    int count = x->c->count[l] + x->c->count[r];
    x->c->count[node] = count;
    float area = box_area(mn, mx);
    x->cost[node] = x->ci * area + x->cost[l] + x->cost[r];
    float asLeaf = x->ct * area * (float)count;
    if (asLeaf < x->cost[node]) {
        x->cost[node] = asLeaf;
        x->flag[node] = 1;
    }
    x->c->data[node] = -1;
    x->deliver[node] = (uint8_t)(x->deliver[l] && x->deliver[r] && x->flag[node]);
}

/* writeSubtreeLeavesNonrecursive (TreeCollapser.cu:95-137): the stack lives
 * at the end of the output buffer; right-first DFS; duplicates removed. */
static int write_leaves(const orc_tree* t, int index, int32_t* leaves, int numberOfLeaves)
{
    uint32_t n = t->n;
    int position = 0;
    leaves[numberOfLeaves - 1] = t->left[index];
    leaves[numberOfLeaves - 2] = t->right[index];
    int stackSize = 2;
    while (stackSize > 0) {
        index = leaves[numberOfLeaves - stackSize];
        --stackSize;
        if (is_leaf(index, n)) {
            int dataIndex = t->data[index];
            int unique = 1;
            for (int i = 0; i < position; ++i) {
                if (leaves[i] == dataIndex) {
                    unique = 0;
                    break;
                }
            }
            if (unique) leaves[position++] = dataIndex;
        } else {
            int left = t->left[index];
            int right = t->right[index];
            leaves[numberOfLeaves - 1 - stackSize++] = left;
            leaves[numberOfLeaves - 1 - stackSize++] = right;
        }
    }
    return position;
}

int orc_collapse(const orc_tree* t, float ci, float ct, orc_collapsed* c, float* sah)
{
    uint32_t n = t->n;
    size_t nodes = 2 * (size_t)n - 1;
    c->n = n;
    c->root = 0;
    for (size_t i = 0; i < nodes; ++i) {
        memcpy(c->bbmax + i * 4, t->bbmax + i * 4, sizeof(float) * 4);
        memcpy(c->bbmin + i * 4, t->bbmin + i * 4, sizeof(float) * 4);
        c->data[i] = t->data[i];
        c->left[i] = t->left[i];
        c->right[i] = t->right[i];
        c->parent[i] = t->parent[i];
        c->count[i] = 0;
    }
    cost_ctx x;
    x.t = t;
    x.c = c;
    x.ci = ci;
    x.ct = ct;
    x.cost = (float*)malloc(sizeof(float) * nodes);
    x.flag = (uint8_t*)calloc(nodes, 1);
    x.deliver = (uint8_t*)calloc(nodes, 1);
    if (!x.cost || !x.flag || !x.deliver) {
        free(x.cost);
        free(x.flag);
        free(x.deliver);
        return -1;
    }
    postorder((orc_tree*)t, &x, cost_visit);

    uint32_t position = 0;
    for (uint32_t p = 0; p + 1 < n; ++p) {
        int l = t->left[p], r = t->right[p];
        if (!(x.deliver[l] && x.deliver[r]) || x.flag[p]) continue;
        int kids[2] = {l, r};
        for (int k = 0; k < 2; ++k) {
            int child = kids[k];
            int cnt = c->count[child];
            if (cnt > 1) {
                c->data[child] = (int32_t)position;
                c->count[child] = write_leaves(t, child, c->tri_index + position, cnt);
                position += (uint32_t)cnt;
            }
        }
    }
    c->used = position;
    if (sah) *sah = x.cost[t->root] / t->area[t->root];
    free(x.cost);
    free(x.flag);
    free(x.deliver);
    return 0;
}

/* --------------------------------------------------------------------- */
/* Intersect::RayTriangle (gpu-ray-traversal/src/rt/Util.cpp:50-94):
 * two-sided Moller-Trumbore, EPSILON = 0.  dot(a,b) = (ax*bx + ay*by) + az*bz. */
int orc_ray_triangle(const float* v0, const float* v1, const float* v2, const float* o,
                     const float* d, float tmin, float tmax, float* t_out, float* u_out,
                     float* v_out)
{
    float e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    float e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    float p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2],
                  d[0] * e2[1] - d[1] * e2[0]};
    float det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    float tv[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
    float u = tv[0] * p[0] + tv[1] * p[1] + tv[2] * p[2];
    float q[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2],
                  tv[0] * e1[1] - tv[1] * e1[0]};
    float v = d[0] * q[0] + d[1] * q[1] + d[2] * q[2];
    if (det > 0.0f) {
        if (u < 0.0f || u > det) return 0;
        if (v < 0.0f || u + v > det) return 0;
    } else if (det < 0.0f) {
        if (u > 0.0f || u < det) return 0;
        if (v > 0.0f || u + v < det) return 0;
    } else {
        return 0;
    }
    float inv = 1.0f / det;
    float t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
    u *= inv;
    v *= inv;
    if (t > tmin && t < tmax) {
        *t_out = t;
        *u_out = u;
        *v_out = v;
        return 1;
    }
    return 0;
}

static int32_t fbits(float f)
{
    int32_t b;
    memcpy(&b, &f, 4);
    return b;
}

/* Conservative slab test.  Direction components below 2^-80 in magnitude are
 * replaced by +-2^-80 before inversion, as the reference GPU kernel does
 * (kepler_dynamic_fetch.cu:128-134), so no 0*inf NaN appears; the entry and
 * exit distances are then widened by 2^-16 relative so that rounding in the
 * slab test never culls a box that holds a Moller-Trumbore hit (the reference
 * CPU tracer divides per axis, Util.cpp:34-46). */
#define ORC_BOX_EPS 1.52587890625e-05f

static int ray_box(const float* mn, const float* mx, const float* o, const float* inv,
                   float tmin, float tmax, float* tnear)
{
    float lo = tmin, hi = tmax;
    for (int k = 0; k < 3; ++k) {
        float t0 = (mn[k] - o[k]) * inv[k];
        float t1 = (mx[k] - o[k]) * inv[k];
        lo = fmaxf(lo, fminf(t0, t1));
        hi = fminf(hi, fmaxf(t0, t1));
    }
    *tnear = lo;
    return lo - fabsf(lo) * ORC_BOX_EPS <= hi + fabsf(hi) * ORC_BOX_EPS;
}

void orc_ray_inverse(const float* d, float* inv)
{
    const float ooeps = 8.271806125530277e-25f; /* 2^-80 */
    for (int k = 0; k < 3; ++k) {
        float dk = fabsf(d[k]) > ooeps ? d[k] : copysignf(ooeps, d[k]);
        inv[k] = 1.0f / dk;
    }
}

typedef struct {
    const orc_collapsed* c;
    const float* v4;
    const float* o;
    const float* d;
    float inv[3];
    float tmin, tmax;
    int any;
    int id;
    float t, u, v;
    int done;
} trace_ctx;

static void leaf_test(trace_ctx* x, int tri)
{
    const float* a = x->v4 + (size_t)tri * 12;
    float t, u, v;
    if (orc_ray_triangle(a, a + 4, a + 8, x->o, x->d, x->tmin, x->tmax, &t, &u, &v)) {
        x->tmax = t;
        x->id = tri;
        x->t = t;
        x->u = u;
        x->v = v;
        if (x->any) x->done = 1;
    }
}

/* BVH::traceRecursive (gpu-ray-traversal/src/rt/bvh/BVH.cpp:89-163) over the
 * collapsed tree: a child is a leaf iff its data index is >= 0
 * (CudaBVH.cpp:504); both children tested, nearer entry visited first.  A
 * collapsed internal node holds a list of count triangles at its data index,
 * a tree leaf its triangle (read by node kind; see orc_compact). */
static void trace_rec(trace_ctx* x, int node)
{
    const orc_collapsed* c = x->c;
    if (x->done) return;
    if (c->data[node] >= 0) {
        int cnt = c->count[node];
        if (node < (int)c->n - 1) {
            for (int j = 0; j < cnt && !x->done; ++j) leaf_test(x, c->tri_index[c->data[node] + j]);
        } else {
            leaf_test(x, c->data[node]);
        }
        return;
    }
    int c0 = c->left[node], c1 = c->right[node];
    float n0, n1;
    int h0 = ray_box(c->bbmin + (size_t)c0 * 4, c->bbmax + (size_t)c0 * 4, x->o, x->inv, x->tmin,
                     x->tmax, &n0);
    int h1 = ray_box(c->bbmin + (size_t)c1 * 4, c->bbmax + (size_t)c1 * 4, x->o, x->inv, x->tmin,
                     x->tmax, &n1);
    if (h0 && h1 && n0 > n1) {
        int tmp = c0;
        c0 = c1;
        c1 = tmp;
    }
    if (h0 && h1) {
        trace_rec(x, c0);
        if (x->done) return;
        trace_rec(x, c1);
    } else if (h0) {
        trace_rec(x, c0);
    } else if (h1) {
        trace_rec(x, c1);
    }
}

void orc_trace(const orc_collapsed* c, const float* v4, const float* rays, uint32_t r,
               int any_hit, int32_t* results)
{
    for (uint32_t i = 0; i < r; ++i) {
        const float* ray = rays + (size_t)i * 8;
        trace_ctx x;
        x.c = c;
        x.v4 = v4;
        x.o = ray;
        x.d = ray + 4;
        orc_ray_inverse(x.d, x.inv);
        x.tmin = ray[3];
        x.tmax = ray[7];
        x.any = any_hit;
        x.id = -1;
        x.t = ray[7];
        x.u = 0.0f;
        x.v = 0.0f;
        x.done = 0;
        if (x.tmax > x.tmin) trace_rec(&x, c->root);
        results[i * 4 + 0] = x.id;
        results[i * 4 + 1] = fbits(x.t);
        results[i * 4 + 2] = fbits(x.u);
        results[i * 4 + 3] = fbits(x.v);
    }
}

void orc_trace_brute(const float* v4, uint32_t n, const float* rays, uint32_t r, int32_t* results)
{
    for (uint32_t i = 0; i < r; ++i) {
        const float* ray = rays + (size_t)i * 8;
        float tmax = ray[7], best_u = 0.0f, best_v = 0.0f;
        int id = -1;
        for (uint32_t k = 0; k < n; ++k) {
            const float* a = v4 + (size_t)k * 12;
            float t, u, v;
            if (orc_ray_triangle(a, a + 4, a + 8, ray, ray + 4, ray[3], tmax, &t, &u, &v)) {
                tmax = t;
                id = (int)k;
                best_u = u;
                best_v = v;
            }
        }
        results[i * 4 + 0] = id;
        results[i * 4 + 1] = fbits(tmax);
        results[i * 4 + 2] = fbits(best_u);
        results[i * 4 + 3] = fbits(best_v);
    }
}

/* --- ray generation ------------------------------------------------------ */
/* jenkinsMix (gpu-ray-traversal/src/rt/ray/RayGenKernels.cu:34-45): Bob
 * Jenkins' 96-bit mix. */
static void orc_jenkins_mix(uint32_t* a, uint32_t* b, uint32_t* c)
{
    *a -= *b; *a -= *c; *a ^= (*c >> 13);
    *b -= *c; *b -= *a; *b ^= (*a << 8);
    *c -= *a; *c -= *b; *c ^= (*b >> 13);
    *a -= *b; *a -= *c; *a ^= (*c >> 12);
    *b -= *c; *b -= *a; *b ^= (*a << 16);
    *c -= *a; *c -= *b; *c ^= (*b >> 5);
    *a -= *b; *a -= *c; *a ^= (*c >> 3);
    *b -= *c; *b -= *a; *b ^= (*a << 10);
    *c -= *a; *c -= *b; *c ^= (*b >> 15);
}

/* normalize() of framework/base/Math.hpp:139-141: v * (1 * rcp(length)),
 * length = sqrt((x*x + y*y) + z*z), rcp(0) = 0 (Math.hpp:112). */
static void orc_normalize3(float* v)
{
    const float l = sqrtf((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    const float r = l != 0.0f ? 1.0f / l : 0.0f;
    v[0] *= r;
    v[1] *= r;
    v[2] *= r;
}

/* rayGenAOKernel (RayGenKernels.cu:114-221) with numSamples = 1, one output
 * ray per input ray, normals as Scene.cpp:74 computes them
 * (normalize(cross(v1 - v0, v2 - v0))).  cosf/sinf are the host C library's;
 * everything else is the kernel's fp32 arithmetic in its order. */
void orc_rays_diffuse(const float* v4, const float* prim, const int32_t* res, uint32_t r,
                      uint32_t seed, float tmax, float* out)
{
    const float pi = 3.14159265358979323846f; /* FW_PI, Defs.hpp:121 */
    for (uint32_t i = 0; i < r; ++i) {
        const float* o = prim + 8ull * i;
        const float* d = o + 4;
        float ht;
        memcpy(&ht, &res[4ull * i + 1], 4);
        const float back = fmaxf(ht - 1.0e-4f, 0.0f);
        const float org[3] = {o[0] + d[0] * back, o[1] + d[1] * back, o[2] + d[2] * back};
        const int tri = res[4ull * i];
        float n[3] = {1.0f, 0.0f, 0.0f};
        if (tri != -1) {
            const float* a = v4 + 12ull * tri;
            const float* b = a + 4;
            const float* c = a + 8;
            const float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
            const float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
            n[0] = e1[1] * e2[2] - e1[2] * e2[1];
            n[1] = e1[2] * e2[0] - e1[0] * e2[2];
            n[2] = e1[0] * e2[1] - e1[1] * e2[0];
            orc_normalize3(n);
        }
        if ((n[0] * d[0] + n[1] * d[1]) + n[2] * d[2] > 0.0f) {
            n[0] = -n[0];
            n[1] = -n[1];
            n[2] = -n[2];
        }
        const float na[3] = {fabsf(n[0]), fabsf(n[1]), fabsf(n[2])};
        const float nm = fmaxf(fmaxf(na[0], na[1]), na[2]);
        float p[3] = {n[1], -n[0], 0.0f};
        if (nm == na[2]) {
            p[0] = 0.0f;
            p[1] = n[2];
            p[2] = -n[1];
        } else if (nm == na[0]) {
            p[0] = -n[2];
            p[1] = 0.0f;
            p[2] = n[0];
        }
        orc_normalize3(p);
        const float bp[3] = {n[1] * p[2] - n[2] * p[1], n[2] * p[0] - n[0] * p[2],
                             n[0] * p[1] - n[1] * p[0]};
        uint32_t ha = seed + i, hb = 0x9e3779b9u, hc = 0x9e3779b9u;
        orc_jenkins_mix(&ha, &hb, &hc);
        orc_jenkins_mix(&ha, &hb, &hc);
        const float angle = 2.0f * pi * (float)hc * exp2f(-32.0f);
        const float ca = cosf(angle), sa = sinf(angle);
        float t0[3], t1[3];
        for (int k = 0; k < 3; ++k) {
            t0[k] = p[k] * ca + bp[k] * sa;
            t1[k] = p[k] * -sa + bp[k] * ca;
        }
        /* sample 0: Halton base 2 of index 1 and base 3 of index 1 */
        float x = 0.0f, xadd = 1.0f;
        for (unsigned int h2 = 1; h2 != 0; h2 >>= 1) {
            xadd *= 0.5f;
            if (h2 & 1) x += xadd;
        }
        float y = 0.0f, yadd = 1.0f;
        for (int h3 = 1; h3 != 0; h3 /= 3) {
            yadd *= 1.0f / 3.0f;
            y += (float)(h3 % 3) * yadd;
        }
        const float phi = 2.0f * pi * y;
        const float rr = sqrtf(x);
        x = rr * cosf(phi);
        y = rr * sinf(phi);
        const float z = sqrtf(1.0f - x * x - y * y);
        float dir[3];
        for (int k = 0; k < 3; ++k) dir[k] = (x * t0[k] + y * t1[k]) + z * n[k];
        orc_normalize3(dir);
        float* q = out + 8ull * i;
        q[0] = org[0];
        q[1] = org[1];
        q[2] = org[2];
        q[3] = 0.0f;
        q[4] = dir[0];
        q[5] = dir[1];
        q[6] = dir[2];
        q[7] = tri == -1 ? -1.0f : tmax;
    }
}

/* --- the reference tracer's Compact layout and Woop test ------------------ */
/* 3x3 determinant of framework/base/Math.hpp:930-935 (detImpl, L = 3), in its
 * operand order. */
static float orc_det3(float m[3][3])
{
    return m[0][0] * m[1][1] * m[2][2] - m[0][0] * m[1][2] * m[2][1] +
           m[1][0] * m[2][1] * m[0][2] - m[1][0] * m[2][2] * m[0][1] +
           m[2][0] * m[0][1] * m[1][2] - m[2][0] * m[0][2] * m[1][1];
}

/* CudaBVH::woopifyTri (gpu-ray-traversal/src/rt/cuda/CudaBVH.cpp:579-600):
 * the inverse of [v0 - v2 | v1 - v2 | (v0 - v2) x (v1 - v2) | v2] (last row
 * 0 0 0 1) by MatrixBase::inverted (Math.hpp:961-980: cofactors, the
 * determinant accumulated over every (i, j), result * rcp(d) * 4); the three
 * rows kept are (row 2 with its last entry negated, row 0, row 1). */
void orc_woopify(const float* v0, const float* v1, const float* v2, float out[12])
{
    float a[3] = {v0[0] - v2[0], v0[1] - v2[1], v0[2] - v2[2]};
    float b[3] = {v1[0] - v2[0], v1[1] - v2[1], v1[2] - v2[2]};
    float c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    float m[4][4];
    for (int r = 0; r < 3; ++r) {
        m[r][0] = a[r];
        m[r][1] = b[r];
        m[r][2] = c[r];
        m[r][3] = v2[r];
    }
    m[3][0] = 0.0f;
    m[3][1] = 0.0f;
    m[3][2] = 0.0f;
    m[3][3] = 1.0f;
    float inv[4][4];
    float d = 0.0f;
    float si = 1.0f;
    for (int i = 0; i < 4; ++i) {
        float sj = si;
        for (int j = 0; j < 4; ++j) {
            float sub[3][3];
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) sub[k][l] = m[k < j ? k : k + 1][l < i ? l : l + 1];
            float dd = orc_det3(sub) * sj;
            inv[i][j] = dd;
            d += dd * m[j][i];
            sj = -sj;
        }
        si = -si;
    }
    const float rd = d != 0.0f ? 1.0f / d : 0.0f;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) inv[i][j] = inv[i][j] * rd * 4.0f;
    out[0] = inv[2][0];
    out[1] = inv[2][1];
    out[2] = inv[2][2];
    out[3] = -inv[2][3];
    for (int k = 0; k < 4; ++k) {
        out[4 + k] = inv[0][k];
        out[8 + k] = inv[1][k];
    }
}

/* CudaBVH::createCompact(const BVHCollapsedTree*, Scene*, nodeOffsetSizeDiv)
 * (CudaBVH.cpp:469-574): a host DFS from the root with an explicit stack
 * (last in, first out); each internal node gets 4 int4 (child boxes' x/y,
 * both z ranges, the two child pointers: an internal child's byte offset /
 * div, ~(first triangle row) for a leaf).  A child is a leaf iff its data
 * index is >= 0 (:504).  A leaf appends 3 Woop rows per triangle (its list,
 * or its single triangle) and a row of 0x80000000 words; tri_index gets the
 * triangle id and two zeros per triangle, 0 per terminator.  With NULL
 * outputs only the sizes are returned (rows of 16 bytes each). */
int orc_compact(const orc_collapsed* c, const float* v4, int div, int32_t* nodes,
                uint32_t* num_nodes, float* woop, int32_t* tri_index, uint32_t* num_rows)
{
    const uint32_t N = c->n;
    if (N < 2 || (div != 1 && div != 16)) return -1;
    int32_t* stack_node = (int32_t*)malloc(sizeof(int32_t) * 2 * N);
    int32_t* stack_slot = (int32_t*)malloc(sizeof(int32_t) * 2 * N);
    if (!stack_node || !stack_slot) {
        free(stack_node);
        free(stack_slot);
        return -1;
    }
    uint32_t nn = 1, rows = 0;
    int sp = 0;
    stack_node[sp] = c->root;
    stack_slot[sp] = 0;
    ++sp;
    while (sp > 0) {
        --sp;
        const int index = stack_node[sp];
        const int slot = stack_slot[sp];
        const int ch[2] = {c->left[index], c->right[index]};
        int32_t cidx[2];
        for (int k = 0; k < 2; ++k) {
            const int child = ch[k];
            if (c->data[child] < 0) {
                cidx[k] = (int32_t)((uint64_t)nn * 64u / (uint32_t)div);
                stack_node[sp] = child;
                stack_slot[sp] = (int32_t)nn;
                ++sp;
                ++nn;
            } else {
                cidx[k] = ~(int32_t)rows;
                /* the reference tests triangleCount > 1 (CudaBVH.cpp:514);
                 * by node kind here, which differs only for a list that
                 * duplicate removal reduced to one triangle */
                const int list = child < (int)c->n - 1;
                const int cnt = list ? c->count[child] : 1;
                for (int j = 0; j < cnt; ++j) {
                    const int tri = list ? c->tri_index[c->data[child] + j] : c->data[child];
                    if (woop) {
                        const float* a = v4 + (size_t)tri * 12;
                        orc_woopify(a, a + 4, a + 8, woop + (size_t)rows * 4);
                        tri_index[rows] = tri;
                        tri_index[rows + 1] = 0;
                        tri_index[rows + 2] = 0;
                    }
                    rows += 3;
                }
                if (woop) {
                    const uint32_t term = 0x80000000u;
                    for (int k2 = 0; k2 < 4; ++k2) memcpy(woop + (size_t)rows * 4 + k2, &term, 4);
                    tri_index[rows] = 0;
                }
                rows += 1;
            }
        }
        if (nodes) {
            const float* lmn = c->bbmin + (size_t)ch[0] * 4;
            const float* lmx = c->bbmax + (size_t)ch[0] * 4;
            const float* rmn = c->bbmin + (size_t)ch[1] * 4;
            const float* rmx = c->bbmax + (size_t)ch[1] * 4;
            int32_t* dst = nodes + (size_t)slot * 16;
            dst[0] = fbits(lmn[0]);
            dst[1] = fbits(lmx[0]);
            dst[2] = fbits(lmn[1]);
            dst[3] = fbits(lmx[1]);
            dst[4] = fbits(rmn[0]);
            dst[5] = fbits(rmx[0]);
            dst[6] = fbits(rmn[1]);
            dst[7] = fbits(rmx[1]);
            dst[8] = fbits(lmn[2]);
            dst[9] = fbits(lmx[2]);
            dst[10] = fbits(rmn[2]);
            dst[11] = fbits(rmx[2]);
            dst[12] = cidx[0];
            dst[13] = cidx[1];
            dst[14] = 0;
            dst[15] = 0;
        }
    }
    free(stack_node);
    free(stack_slot);
    if (num_nodes) *num_nodes = nn;
    if (num_rows) *num_rows = rows;
    return 0;
}

/* Integer min/max of float bit patterns, as the reference tracer's video
 * instructions compare them (CudaTracerKernels.hpp:165-184). */
static float orc_imin(float a, float b) { return fbits(a) < fbits(b) ? a : b; }
static float orc_imax(float a, float b) { return fbits(a) > fbits(b) ? a : b; }

/* The closest-hit / any-hit result of kepler_dynamic_fetch's `trace`
 * (gpu-ray-traversal/src/rt/kernels/kepler_dynamic_fetch.cu:61-320) on the
 * Compact2 layout (node pointers in 16-byte rows), for one ray at a time:
 * slab spans of spanBeginKepler / spanEndKepler
 * (CudaTracerKernels.hpp:209-210), the nearer child first (a tie keeps child
 * 0), leaves in the order the traversal reaches them, and the Woop test
 * (:257-277) with tmin < t < hitT, u >= 0, v >= 0, u + v <= 1.  Deferring a
 * leaf, as the kernel does, only visits extra nodes with a staler hitT and
 * leaves the result unchanged.  results int32[4r]: {id or -1, t bits, 0, 0}
 * where id is tri_index[row of the hit]. */
void orc_trace_compact(const int32_t* nodes, const float* woop, const int32_t* tri_index,
                       const float* rays, uint32_t r, int any_hit, int32_t* results)
{
    const int32_t sentinel = 0x76543210;
    int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * 4096);
    if (!stack) return;
    for (uint32_t i = 0; i < r; ++i) {
        const float* o = rays + 8ull * i;
        const float* d = o + 4;
        const float ooeps = exp2f(-80.0f);
        const float idx = 1.0f / (fabsf(d[0]) > ooeps ? d[0] : copysignf(ooeps, d[0]));
        const float idy = 1.0f / (fabsf(d[1]) > ooeps ? d[1] : copysignf(ooeps, d[1]));
        const float idz = 1.0f / (fabsf(d[2]) > ooeps ? d[2] : copysignf(ooeps, d[2]));
        const float ox = o[0] * idx, oy = o[1] * idy, oz = o[2] * idz;
        const float tmin = o[3];
        float hitT = d[3];
        int hit = -1;
        int sp = 0;
        stack[sp] = sentinel;
        int32_t node = 0;
        while (node != sentinel) {
            if (node >= 0) {
                const int32_t* q = nodes + (size_t)node * 4;
                float n[12];
                memcpy(n, q, sizeof(n));
                const float c0lox = n[0] * idx - ox, c0hix = n[1] * idx - ox;
                const float c0loy = n[2] * idy - oy, c0hiy = n[3] * idy - oy;
                const float c0loz = n[8] * idz - oz, c0hiz = n[9] * idz - oz;
                const float c1loz = n[10] * idz - oz, c1hiz = n[11] * idz - oz;
                const float c1lox = n[4] * idx - ox, c1hix = n[5] * idx - ox;
                const float c1loy = n[6] * idy - oy, c1hiy = n[7] * idy - oy;
                const float c0min = orc_imax(orc_imax(fminf(c0lox, c0hix), fminf(c0loy, c0hiy)),
                                             orc_imax(orc_imin(c0loz, c0hiz), tmin));
                const float c0max = orc_imin(orc_imin(fmaxf(c0lox, c0hix), fmaxf(c0loy, c0hiy)),
                                             orc_imin(orc_imax(c0loz, c0hiz), hitT));
                const float c1min = orc_imax(orc_imax(fminf(c1lox, c1hix), fminf(c1loy, c1hiy)),
                                             orc_imax(orc_imin(c1loz, c1hiz), tmin));
                const float c1max = orc_imin(orc_imin(fmaxf(c1lox, c1hix), fmaxf(c1loy, c1hiy)),
                                             orc_imin(orc_imax(c1loz, c1hiz), hitT));
                const int t0 = c0max >= c0min, t1 = c1max >= c1min;
                int32_t a = q[12], b = q[13];
                if (!t0 && !t1) {
                    node = stack[sp--];
                } else {
                    node = t0 ? a : b;
                    if (t0 && t1) {
                        if (c1min < c0min) {
                            const int32_t tmp = node;
                            node = b;
                            b = tmp;
                        }
                        stack[++sp] = b;
                    }
                }
                continue;
            }
            /* leaf: rows ~node, ~node + 3, ... up to the terminator */
            int stop = 0;
            for (int row = ~node;; row += 3) {
                const float* v00 = woop + (size_t)row * 4;
                const float* v11 = v00 + 4;
                const float* v22 = v00 + 8;
                if (fbits(v00[0]) == (int32_t)0x80000000) break;
                const float Oz = v00[3] - o[0] * v00[0] - o[1] * v00[1] - o[2] * v00[2];
                const float invDz = 1.0f / (d[0] * v00[0] + d[1] * v00[1] + d[2] * v00[2]);
                const float t = Oz * invDz;
                if (t > tmin && t < hitT) {
                    const float Ox = v11[3] + o[0] * v11[0] + o[1] * v11[1] + o[2] * v11[2];
                    const float Dx = d[0] * v11[0] + d[1] * v11[1] + d[2] * v11[2];
                    const float u = Ox + t * Dx;
                    if (u >= 0.0f) {
                        const float Oy = v22[3] + o[0] * v22[0] + o[1] * v22[1] + o[2] * v22[2];
                        const float Dy = d[0] * v22[0] + d[1] * v22[1] + d[2] * v22[2];
                        const float v = Oy + t * Dy;
                        if (v >= 0.0f && u + v <= 1.0f) {
                            hitT = t;
                            hit = row;
                            if (any_hit) {
                                stop = 1;
                                break;
                            }
                        }
                    }
                }
            }
            if (stop) break;
            node = stack[sp--];
        }
        int32_t* out = results + 4ull * i;
        out[0] = hit < 0 ? -1 : tri_index[hit];
        out[1] = fbits(hitT);
        out[2] = 0;
        out[3] = 0;
    }
    free(stack);
}

/* --------------------------------------------------------------------- */
/* Ray coherence sort: RayBuffer::mortonSort (gpu-ray-traversal/src/rt/ray/
 * RayBuffer.cpp:103-163).  findAABBKernel (RayBufferKernels.cu:64-135): box of
 * the origins and end points o + d*tmax (end points that are not finite are
 * left out).  genMortonKeysKernel (:139-171): a = (o - lo)/(hi - lo),
 * b = (normalize(d) + 1)*0.5, quantised by 2^24 and 2^21 with saturating
 * conversion, bit i of coordinate k at key bit k + 6i (collectBits :57-62);
 * IEEE division and square root.  keys: word w of ray i at keys[w*r + i]. */
static uint32_t sat_u32(float x)
{
    if (!(x >= 1.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

void orc_ray_box(const float* rays, uint32_t r, float box[6])
{
    for (int a = 0; a < 3; ++a) {
        box[a] = FLT_MAX;
        box[3 + a] = -FLT_MAX;
    }
    for (uint32_t i = 0; i < r; ++i) {
        const float* o = rays + (size_t)i * 8;
        const float* d = o + 4;
        float e[3];
        int fin = 1;
        for (int a = 0; a < 3; ++a) {
            box[a] = ORC_MIN(box[a], o[a]);
            box[3 + a] = ORC_MAX(box[3 + a], o[a]);
            e[a] = o[a] + d[a] * d[3];
            fin = fin && isfinite(e[a]);
        }
        if (fin)
            for (int a = 0; a < 3; ++a) {
                box[a] = ORC_MIN(box[a], e[a]);
                box[3 + a] = ORC_MAX(box[3 + a], e[a]);
            }
    }
}

void orc_ray_keys(const float* rays, uint32_t r, const float box[6], uint64_t* keys)
{
    for (uint32_t i = 0; i < r; ++i) {
        const float* o = rays + (size_t)i * 8;
        float d[3] = {o[4], o[5], o[6]};
        orc_normalize3(d);
        uint32_t q[6];
        for (int a = 0; a < 3; ++a) {
            const float av = (o[a] - box[a]) / (box[3 + a] - box[a]);
            q[a] = sat_u32(av * 256.0f * 65536.0f);
            const float bv = (d[a] + 1.0f) * 0.5f;
            q[3 + a] = sat_u32(bv * 32.0f * 65536.0f);
        }
        uint64_t w[3] = {0, 0, 0};
        for (int k = 0; k < 6; ++k)
            for (int b = 0; b < 32; ++b) {
                const int pos = k + 6 * b;
                w[pos >> 6] |= (uint64_t)((q[k] >> b) & 1u) << (pos & 63);
            }
        for (int a = 0; a < 3; ++a) keys[(size_t)a * r + i] = w[a];
    }
}
