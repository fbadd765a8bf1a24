// The reference tracer's Compact / Compact2 BVH layout emitted on the GPU
// from a collapsed tree (CudaBVH::createCompact, gpu-ray-traversal/src/rt/
// cuda/CudaBVH.cpp:469-574, with woopifyTri :579-600), byte for byte the
// layout the reference's host DFS writes, and a traversal kernel with the
// reference's `trace` argument list (rt/kernels/CudaTracerKernels.hpp:73-86)
// and kepler_dynamic_fetch semantics (rt/kernels/kepler_dynamic_fetch.cu:
// 61-320): Kepler slab spans, nearer child first, Woop triangle test.
//
// The reference visits internal nodes with a last-in-first-out stack from the
// root; an internal node's children receive node slots in discovery order
// (left, then right) when the node is processed, and leaf children append
// their Woop rows then.  Processing order is therefore the preorder that
// visits the right subtree first.  On the GPU:
//   1. bottom-up: S(x) = number of emitted (internal, not collapsed) nodes in
//      x's subtree, through per-node arrival counters;
//   2. each emitted node climbs to the root and sums its preorder position
//      (a right child comes right after its parent, a left child after its
//      parent and the whole right sibling subtree) and checks that no
//      ancestor is a collapsed leaf;
//   3. exclusive scans, in that order, of the internal children per node
//      (slots) and of the leaf rows per node (3 per triangle + a terminator);
//   4. one thread per emitted node writes its 64-byte record and its leaf
//      children's Woop rows.
#include "device.hpp"
#include "internal.hpp"

namespace atr {

namespace {

// ------------------------------------------------------------------ layout
// Subtree counts of emitted nodes, bottom-up from the tree leaves.  Every
// hand-over uses the device-coherent protocol of device.hpp (this kernel runs
// once per layout, outside the build).
__global__ __launch_bounds__(256) void k_compact_count(uint32_t N,
                                                       const int32_t* __restrict__ parent,
                                                       const int32_t* __restrict__ left,
                                                       const int32_t* __restrict__ right,
                                                       const int32_t* __restrict__ data,
                                                       uint32_t* counters, int32_t* S)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= N) return;
    int node = (int)(N - 1 + i);
    int cur = parent[node];
    while (cur >= 0) {
        const uint32_t old = arrive_crossing(&counters[cur]);
        if (old == 0u) return;
        const int l = left[cur], r = right[cur];
        const int sl = l >= (int)N - 1 ? 0 : ld<true>(&S[l]);
        const int sr = r >= (int)N - 1 ? 0 : ld<true>(&S[r]);
        const int s = data[cur] >= 0 ? 0 : 1 + sl + sr;
        st_coh(&S[cur], s);
        node = cur;
        cur = parent[cur];
    }
}

// Preorder position (right subtree first) of every emitted node; -1 for
// internal nodes that are not emitted (collapsed leaves and nodes below them).
__global__ __launch_bounds__(256) void k_compact_pos(uint32_t N, const int32_t* __restrict__ parent,
                                                     const int32_t* __restrict__ right,
                                                     const int32_t* __restrict__ data,
                                                     const int32_t* __restrict__ S,
                                                     int32_t* __restrict__ pos_of,
                                                     int32_t* __restrict__ order)
{
    const uint32_t x = blockIdx.x * 256u + threadIdx.x;
    if (x + 1 >= N) return;
    int pos = -1;
    if (data[x] < 0) {
        int c = (int)x;
        int p = 0;
        bool reachable = true;
        while (c != 0) {
            const int par = parent[c];
            if (data[par] >= 0) {
                reachable = false;
                break;
            }
            const int r = right[par];
            if (c == r) {
                p += 1;
            } else {
                p += 1 + (r >= (int)N - 1 ? 0 : S[r]);
            }
            c = par;
        }
        if (reachable && p < (int)N - 1) {
            pos = p;
            order[p] = (int)x;
        }
    }
    pos_of[x] = pos;
}

// Leaves of a collapsed tree are read by node kind: a tree leaf (index
// >= N-1) holds its triangle in its data index, a collapsed internal node a
// list of count[] triangles at that offset.  The reference's createCompact
// tells them apart by triangleCount > 1 (CudaBVH.cpp:514-533), which agrees
// except for a list that the collapse's duplicate removal (trees built from
// split triangles) reduced to one triangle; that list is read as a list.
__device__ __forceinline__ int leaf_rows(int child, uint32_t N, const int32_t* count)
{
    if (child >= (int)N - 1) return 4;
    return 3 * count[child] + 1;
}

__device__ __forceinline__ bool is_inner(int child, uint32_t N, const int32_t* data)
{
    return child < (int)N - 1 && data[child] < 0;
}

__global__ __launch_bounds__(256) void k_compact_sizes(uint32_t M, uint32_t N,
                                                       const int32_t* __restrict__ order,
                                                       const int32_t* __restrict__ left,
                                                       const int32_t* __restrict__ right,
                                                       const int32_t* __restrict__ data,
                                                       const int32_t* __restrict__ count,
                                                       uint32_t* __restrict__ inner_kids,
                                                       uint32_t* __restrict__ rows,
                                                       uint32_t* err)
{
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= M) return;
    const int x = order[p];
    uint32_t k = 0, r = 0;
    if (x < 0 || x + 1 >= (int)N) {  // unfilled position: inconsistent counts
        inner_kids[p] = 0;
        rows[p] = 0;
        atomicOr(err, 4u);
        return;
    }
    for (int side = 0; side < 2; ++side) {
        const int ch = side == 0 ? left[x] : right[x];
        if (is_inner(ch, N, data)) ++k;
        else r += (uint32_t)leaf_rows(ch, N, count);
    }
    inner_kids[p] = k;
    rows[p] = r;
}

// 3x3 determinant in the operand order of framework/base/Math.hpp:930-935.
__device__ __forceinline__ float det3(const float m[3][3])
{
    return m[0][0] * m[1][1] * m[2][2] - m[0][0] * m[1][2] * m[2][1] +
           m[1][0] * m[2][1] * m[0][2] - m[1][0] * m[2][2] * m[0][1] +
           m[2][0] * m[0][1] * m[1][2] - m[2][0] * m[0][2] * m[1][1];
}

// woopifyTri (CudaBVH.cpp:579-600): rows 2 (last entry negated), 0 and 1 of
// the inverse of [v0-v2 | v1-v2 | (v0-v2)x(v1-v2) | v2], by the cofactor
// inversion of Math.hpp:961-980 (determinant summed over every cofactor,
// result * rcp(d) * 4).
__device__ void woopify(const float4 v0, const float4 v1, const float4 v2, float4* out)
{
    const float a[3] = {v0.x - v2.x, v0.y - v2.y, v0.z - v2.z};
    const float b[3] = {v1.x - v2.x, v1.y - v2.y, v1.z - v2.z};
    const float c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                        a[0] * b[1] - a[1] * b[0]};
    const float v2a[3] = {v2.x, v2.y, v2.z};
    float m[4][4];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        m[r][0] = a[r];
        m[r][1] = b[r];
        m[r][2] = c[r];
        m[r][3] = v2a[r];
    }
    m[3][0] = 0.0f;
    m[3][1] = 0.0f;
    m[3][2] = 0.0f;
    m[3][3] = 1.0f;
    float inv[4][4];
    float d = 0.0f;
    float si = 1.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float sj = si;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float sub[3][3];
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int l = 0; l < 3; ++l) sub[k][l] = m[k < j ? k : k + 1][l < i ? l : l + 1];
            const float dd = det3(sub) * sj;
            inv[i][j] = dd;
            d += dd * m[j][i];
            sj = -sj;
        }
        si = -si;
    }
    const float rd = d != 0.0f ? 1.0f / d : 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) inv[i][j] = inv[i][j] * rd * 4.0f;
    out[0] = make_float4(inv[2][0], inv[2][1], inv[2][2], -inv[2][3]);
    out[1] = make_float4(inv[0][0], inv[0][1], inv[0][2], inv[0][3]);
    out[2] = make_float4(inv[1][0], inv[1][1], inv[1][2], inv[1][3]);
}

__global__ __launch_bounds__(256) void k_compact_emit(
    uint32_t M, uint32_t R, uint32_t N, int div, const int32_t* __restrict__ order,
    const int32_t* __restrict__ pos_of, const int32_t* __restrict__ parent,
    const int32_t* __restrict__ left, const int32_t* __restrict__ right,
    const int32_t* __restrict__ data, const int32_t* __restrict__ count,
    const int32_t* __restrict__ tri_index, const float4* __restrict__ bbmin,
    const float4* __restrict__ bbmax, const float4* __restrict__ v,
    const uint32_t* __restrict__ kid_base, const uint32_t* __restrict__ row_base,
    int4* __restrict__ nodes, float4* __restrict__ woop, int32_t* __restrict__ rindex,
    uint32_t* err)
{
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= M) return;
    const int x = order[p];
    if (x < 0 || x + 1 >= (int)N) return;  // flagged by k_compact_sizes
    // this node's slot, assigned when its parent was processed
    int slot = 0;
    if (x != 0) {
        const int par = parent[x];
        const int pl = left[par];
        slot = 1 + (int)kid_base[pos_of[par]] + ((x != pl && is_inner(pl, N, data)) ? 1 : 0);
    }
    const int ch[2] = {left[x], right[x]};
    int cidx[2];
    int kslot = 1 + (int)kid_base[p];
    uint32_t row = row_base[p];
    uint32_t need = 0;
    for (int k = 0; k < 2; ++k)
        if (!is_inner(ch[k], N, data)) need += (uint32_t)leaf_rows(ch[k], N, count);
    if (slot < 0 || (uint32_t)slot >= M || (uint64_t)row + need > R) {
        atomicOr(err, 4u);
        return;
    }
    for (int k = 0; k < 2; ++k) {
        const int c = ch[k];
        if (is_inner(c, N, data)) {
            cidx[k] = (int)((uint64_t)kslot * 64u / (uint32_t)div);
            ++kslot;
        } else {
            cidx[k] = ~(int)row;
            const bool list = c < (int)N - 1;
            const int cnt = list ? count[c] : 1;
            for (int j = 0; j < cnt; ++j) {
                const int tri = list ? tri_index[data[c] + j] : data[c];
                woopify(v[3ull * tri], v[3ull * tri + 1], v[3ull * tri + 2], woop + row);
                rindex[row] = tri;
                rindex[row + 1] = 0;
                rindex[row + 2] = 0;
                row += 3;
            }
            const float term = __int_as_float((int)0x80000000);
            woop[row] = make_float4(term, term, term, term);
            rindex[row] = 0;
            row += 1;
        }
    }
    const float4 lmn = bbmin[ch[0]], lmx = bbmax[ch[0]];
    const float4 rmn = bbmin[ch[1]], rmx = bbmax[ch[1]];
    int4* dst = nodes + 4ull * slot;
    dst[0] = make_int4(__float_as_int(lmn.x), __float_as_int(lmx.x), __float_as_int(lmn.y),
                       __float_as_int(lmx.y));
    dst[1] = make_int4(__float_as_int(rmn.x), __float_as_int(rmx.x), __float_as_int(rmn.y),
                       __float_as_int(rmx.y));
    dst[2] = make_int4(__float_as_int(lmn.z), __float_as_int(lmx.z), __float_as_int(rmn.z),
                       __float_as_int(rmx.z));
    dst[3] = make_int4(cidx[0], cidx[1], 0, 0);
}

__global__ void k_compact_fill(uint32_t* a, uint32_t* b, uint64_t count)
{
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < count; i += 256ull * gridDim.x) {
        a[i] = 0u;
        b[i] = 0xffffffffu;
    }
}

// ------------------------------------------------------------------ trace
// Integer min/max of float bit patterns (the reference's vmin/vmax video
// instructions, CudaTracerKernels.hpp:165-184): ordered like the floats for
// non-negative values and for a negative against a non-negative one.
__device__ __forceinline__ float imin(float a, float b)
{
    return __int_as_float(min(__float_as_int(a), __float_as_int(b)));
}
__device__ __forceinline__ float imax(float a, float b)
{
    return __int_as_float(max(__float_as_int(a), __float_as_int(b)));
}

constexpr int kSentinel = 0x76543210;  // CudaTracerKernels.hpp:38
constexpr int kTraceBlock = 256;
constexpr int kRing = 16;              // traversal stack entries per lane kept in LDS

template <bool ANY>
__global__ __launch_bounds__(kTraceBlock) void k_trace_compact(
    int num_rays, const float4* __restrict__ rays, int2* __restrict__ results,
    const float4* __restrict__ nodes, const float4* __restrict__ tris,
    const int32_t* __restrict__ tri_index, uint32_t* __restrict__ counter, uint32_t* err)
{
    // the newest 16 entries per lane in LDS, 112 more in private memory
    __shared__ uint32_t s_ring[kRing * kTraceBlock];
    LdsStack<kRing, 112, kTraceBlock> stack(s_ring + threadIdx.x);
    const uint64_t lt = lanemask_lt();

    bool active = false;
    RayQueue q(counter, (uint32_t)num_rays);
    int ray = 0;
    float ox = 0.0f, oy = 0.0f, oz = 0.0f, tmin = 0.0f, dx = 0.0f, dy = 0.0f, dz = 0.0f;
    float hitT = 0.0f, idx = 0.0f, idy = 0.0f, idz = 0.0f, oodx = 0.0f, oody = 0.0f, oodz = 0.0f;
    int hit = -1, node = kSentinel;
    while (true) {
        // fetch rays for idle lanes once a quarter of the wave is idle
        const uint64_t idle = __ballot(!active);
        if (!q.exhausted && __popcll(idle) >= 16) {
            const int leader = __ffsll((unsigned long long)idle) - 1;
            uint32_t first = 0;
            const uint32_t take = q.take((uint32_t)__popcll(idle), leader, first);
            const uint32_t rank = (uint32_t)__popcll(idle & lt);
            if (!active && rank < take) {
                const uint32_t k = q.ray_of(first + rank);
                if (k < (uint32_t)num_rays) {
                    ray = (int)k;
                    const float4 o = rays[2ull * k];
                    const float4 d = rays[2ull * k + 1];
                    ox = o.x;
                    oy = o.y;
                    oz = o.z;
                    tmin = o.w;
                    dx = d.x;
                    dy = d.y;
                    dz = d.z;
                    hitT = d.w;
                    const float ooeps = 8.271806125530277e-25f;  // exp2f(-80)
                    idx = 1.0f / (fabsf(dx) > ooeps ? dx : copysignf(ooeps, dx));
                    idy = 1.0f / (fabsf(dy) > ooeps ? dy : copysignf(ooeps, dy));
                    idz = 1.0f / (fabsf(dz) > ooeps ? dz : copysignf(ooeps, dz));
                    oodx = ox * idx;
                    oody = oy * idy;
                    oodz = oz * idz;
                    stack.n = 0;
                    node = 0;
                    hit = -1;
                    active = true;
                }
            }
        }
        if (__ballot(active) == 0ull) {
            if (q.exhausted) break;
            continue;
        }
        // internal nodes until every active lane stands at a leaf or is done
        while (__ballot(active && node >= 0 && node != kSentinel) != 0ull) {
            if (active && node >= 0 && node != kSentinel) {
                const float4 n0 = nodes[node + 0];
                const float4 n1 = nodes[node + 1];
                const float4 nz = nodes[node + 2];
                const float4 tc = nodes[node + 3];
                const float c0lox = n0.x * idx - oodx, c0hix = n0.y * idx - oodx;
                const float c0loy = n0.z * idy - oody, c0hiy = n0.w * idy - oody;
                const float c0loz = nz.x * idz - oodz, c0hiz = nz.y * idz - oodz;
                const float c1loz = nz.z * idz - oodz, c1hiz = nz.w * idz - oodz;
                const float c1lox = n1.x * idx - oodx, c1hix = n1.y * idx - oodx;
                const float c1loy = n1.z * idy - oody, c1hiy = n1.w * idy - oody;
                const float c0min = imax(imax(fminf(c0lox, c0hix), fminf(c0loy, c0hiy)),
                                         imax(imin(c0loz, c0hiz), tmin));
                const float c0max = imin(imin(fmaxf(c0lox, c0hix), fmaxf(c0loy, c0hiy)),
                                         imin(imax(c0loz, c0hiz), hitT));
                const float c1min = imax(imax(fminf(c1lox, c1hix), fminf(c1loy, c1hiy)),
                                         imax(imin(c1loz, c1hiz), tmin));
                const float c1max = imin(imin(fmaxf(c1lox, c1hix), fmaxf(c1loy, c1hiy)),
                                         imin(imax(c1loz, c1hiz), hitT));
                const bool t0 = c0max >= c0min, t1 = c1max >= c1min;
                int a = __float_as_int(tc.x), b = __float_as_int(tc.y);
                if (!t0 && !t1) {
                    node = stack.n > 0 ? (int)stack.pop() : kSentinel;
                } else {
                    node = t0 ? a : b;
                    if (t0 && t1) {
                        if (c1min < c0min) {
                            const int tmp = node;
                            node = b;
                            b = tmp;
                        }
                        if (!stack.push((uint32_t)b)) {
                            atomicOr(err, 2u);
                            node = kSentinel;
                            hit = -1;
                        }
                    }
                }
            }
        }
        if (active && node != kSentinel) {
            // leaf: Woop rows ~node, ~node + 3, ... up to the terminator
            bool stop = false;
            for (int row = ~node;; row += 3) {
                const float4 v00 = tris[row];
                if (__float_as_int(v00.x) == (int)0x80000000) break;
                const float4 v11 = tris[row + 1];
                const float4 v22 = tris[row + 2];
                const float Oz = v00.w - ox * v00.x - oy * v00.y - oz * v00.z;
                const float invDz = 1.0f / (dx * v00.x + dy * v00.y + dz * v00.z);
                const float t = Oz * invDz;
                if (t > tmin && t < hitT) {
                    const float Ox = v11.w + ox * v11.x + oy * v11.y + oz * v11.z;
                    const float Dx = dx * v11.x + dy * v11.y + dz * v11.z;
                    const float u = Ox + t * Dx;
                    if (u >= 0.0f) {
                        const float Oy = v22.w + ox * v22.x + oy * v22.y + oz * v22.z;
                        const float Dy = dx * v22.x + dy * v22.y + dz * v22.z;
                        const float vv = Oy + t * Dy;
                        if (vv >= 0.0f && u + vv <= 1.0f) {
                            hitT = t;
                            hit = row;
                            if (ANY) {
                                stop = true;
                                break;
                            }
                        }
                    }
                }
            }
            node = (stop || stack.n == 0) ? kSentinel : (int)stack.pop();
        }
        if (active && node == kSentinel) {
            results[2ull * ray] = make_int2(hit < 0 ? -1 : tri_index[hit], __float_as_int(hitT));
            active = false;
        }
    }
}

__global__ void k_zero_counter(uint32_t* p) { p[threadIdx.x] = 0u; }

}  // namespace

// ------------------------------------------------------------------ host
atr_status run_compact(atr_context* ctx, atr_collapsed* c, const float4* verts, int div)
{
    const uint32_t N = c->n;
    const uint32_t I = N - 1;  // internal node slots
    int32_t* S = nullptr;
    int32_t* pos_of = nullptr;
    int32_t* order = nullptr;
    uint32_t* counters = nullptr;
    uint32_t* kids = nullptr;
    uint32_t* rows = nullptr;
    uint32_t* kid_base = nullptr;
    uint32_t* row_base = nullptr;
    uint32_t* totals = nullptr;
    atr_status s = ATR_OK;
    if (s == ATR_OK) s = dalloc_t(ctx, I, &S);
    if (s == ATR_OK) s = dalloc_t(ctx, I, &pos_of);
    if (s == ATR_OK) s = dalloc_t(ctx, I, &order);
    if (s == ATR_OK) s = dalloc_t(ctx, I, &counters);
    if (s == ATR_OK) s = dalloc_t(ctx, I, &kids);
    if (s == ATR_OK) s = dalloc_t(ctx, I, &rows);
    if (s == ATR_OK) s = dalloc_t(ctx, I, &kid_base);
    if (s == ATR_OK) s = dalloc_t(ctx, I, &row_base);
    if (s == ATR_OK) s = dalloc_t(ctx, 4, &totals);
    uint32_t M = 0, R = 0;
    if (s == ATR_OK)
        s = launch(ctx, "compact_zero", I * 4.0, [&] {
            k_compact_fill<<<std::min<uint32_t>(div_up(I, 256), 2048u), 256, 0, ctx->stream>>>(
                counters, (uint32_t*)order, I);
        });
    if (s == ATR_OK)
        s = launch(ctx, "compact_count", 0.0, [&] {
            k_compact_count<<<div_up(N, 256), 256, 0, ctx->stream>>>(N, c->parent, c->left,
                                                                      c->right, c->data, counters, S);
        });
    if (s == ATR_OK) {
        // the root's count is the number of emitted nodes
        hipError_t e = hipMemcpyAsync(&M, S, 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) s = fail(ATR_ERR_DEVICE, "compact node count readback failed");
        if (s == ATR_OK && (M == 0 || M > I)) s = fail(ATR_ERR_INTERNAL, "bad emitted node count");
    }
    if (s == ATR_OK)
        s = launch(ctx, "compact_pos", 0.0, [&] {
            k_compact_pos<<<div_up(I, 256), 256, 0, ctx->stream>>>(N, c->parent, c->right, c->data,
                                                                    S, pos_of, order);
        });
    if (s == ATR_OK)
        s = launch(ctx, "compact_sizes", 0.0, [&] {
            k_compact_sizes<<<div_up(M, 256), 256, 0, ctx->stream>>>(
                M, N, order, c->left, c->right, c->data, c->count, kids, rows, ctx->d_status);
        });
    if (s == ATR_OK) s = run_scan_u32(ctx, kids, M, kid_base, totals);
    if (s == ATR_OK) s = run_scan_u32(ctx, rows, M, row_base, totals + 1);
    uint32_t tot[2] = {0, 0};
    if (s == ATR_OK) {
        hipError_t e = hipMemcpyAsync(tot, totals, 8, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) s = fail(ATR_ERR_DEVICE, "compact size readback failed");
        R = tot[1];
        if (s == ATR_OK && tot[0] + 1 != M) s = fail(ATR_ERR_INTERNAL, "compact slot count mismatch");
    }
    if (s == ATR_OK) {
        dfree(ctx, c->cnodes);
        dfree(ctx, c->cwoop);
        dfree(ctx, c->crows_index);
        c->cnodes = nullptr;
        c->cwoop = nullptr;
        c->crows_index = nullptr;
        s = dalloc_t(ctx, 4ull * M, &c->cnodes);
        if (s == ATR_OK) s = dalloc_t(ctx, R, &c->cwoop);
        if (s == ATR_OK) s = dalloc_t(ctx, R, &c->crows_index);
    }
    if (s == ATR_OK)
        s = launch(ctx, "compact_emit", (double)M * 64.0 + (double)R * 20.0, [&] {
            k_compact_emit<<<div_up(M, 256), 256, 0, ctx->stream>>>(
                M, R, N, div, order, pos_of, c->parent, c->left, c->right, c->data, c->count,
                c->tri_index, c->bbmin, c->bbmax, verts, kid_base, row_base, (int4*)c->cnodes,
                c->cwoop, c->crows_index, ctx->d_status);
        });
    if (s == ATR_OK) {
        c->compact_nodes = M;
        c->compact_rows = R;
        c->compact_div = div;
    }
    for (void* p : {(void*)S, (void*)pos_of, (void*)order, (void*)counters, (void*)kids,
                    (void*)rows, (void*)kid_base, (void*)row_base, (void*)totals})
        dfree(ctx, p);
    return s;
}

atr_status run_trace_compact(atr_context* ctx, int num_rays, int any_hit, const float4* rays,
                             int2* results, const float4* nodes, const float4* tris,
                             const int32_t* tri_index)
{
    if (num_rays <= 0) return ATR_OK;
    if (ctx->cu_count == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) !=
                hipSuccess ||
            cus <= 0)
            cus = 256;
        ctx->cu_count = cus;
    }
    const uint32_t grid = std::min<uint32_t>(div_up((uint32_t)num_rays, kTraceBlock),
                                             (uint32_t)ctx->cu_count * 8u);
    uint32_t* counter = nullptr;
    ATR_TRY(dalloc_t(ctx, kRayCounterWords, &counter));
    atr_status s = launch(ctx, "trace_compact_zero", kRayCounterWords * 4.0, [&] {
        k_zero_counter<<<1, kRayCounterWords, 0, ctx->stream>>>(counter);
    });
    if (s == ATR_OK)
        s = launch(ctx, any_hit ? "trace_compact_any" : "trace_compact", num_rays * 40.0, [&] {
            if (any_hit)
                k_trace_compact<true><<<grid, kTraceBlock, 0, ctx->stream>>>(
                    num_rays, rays, results, nodes, tris, tri_index, counter, ctx->d_status);
            else
                k_trace_compact<false><<<grid, kTraceBlock, 0, ctx->stream>>>(
                    num_rays, rays, results, nodes, tris, tri_index, counter, ctx->d_status);
        });
    dfree(ctx, counter);
    return s;
}

}  // namespace atr
