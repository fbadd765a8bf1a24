// LBVH on gfx950: scene box, Morton encode (+ fused per-block digit
// histograms), onesweep LSD radix sort of (key, index), and a fused
// bottom-up radix-tree build + AABB refit that reproduces the Karras node
// numbering of BVHRT-Core/src/LBVH.cu:208-316 exactly.
#include "device.hpp"
#include "internal.hpp"

namespace atr {

namespace {

constexpr int kEncBlock = 256;
constexpr int kEncItems = 16;
constexpr int kEncTile = kEncBlock * kEncItems;  // triangles per encode block

// Radix-sort tiles hold 8192 keys: 1024 threads x 8 keys (the default) or
// 512 threads x 16.  At 2M keys that is one tile per CU, all resident at
// once; against 4096-key tiles of 256 threads a 63-bit pass takes 28 us
// instead of 37 (a shorter look-back chain, half the tiles to dispatch and
// more waves per tile to hide the loads).
constexpr int kSortTile = 8192;

constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagInc = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1u;

// ---------------------------------------------------------------- scene box
// Orderable unsigned image of a float: unsigned order equals float order.
__device__ __forceinline__ uint32_t f2o(float f)
{
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t o)
{
    uint32_t b = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    return __uint_as_float(b);
}

__global__ void k_bbox_init(uint32_t* acc)
{
    const int t = threadIdx.x;
    // minimum starts at FLT_MAX, maximum at FLT_MIN (SceneConverter.cpp:28-34)
    if (t < 3) acc[t] = f2o(FLT_MAX);
    else if (t < 6) acc[t] = f2o(FLT_MIN);
}

__global__ __launch_bounds__(256) void k_bbox(const float4* __restrict__ v, uint64_t count,
                                              uint32_t* acc)
{
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float mx[3] = {FLT_MIN, FLT_MIN, FLT_MIN};
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < count; i += 256ull * gridDim.x) {
        float4 p = v[i];
        mn[0] = fminf(mn[0], p.x);
        mn[1] = fminf(mn[1], p.y);
        mn[2] = fminf(mn[2], p.z);
        mx[0] = fmaxf(mx[0], p.x);
        mx[1] = fmaxf(mx[1], p.y);
        mx[2] = fmaxf(mx[2], p.z);
    }
    __shared__ uint32_t s[6];
    if (threadIdx.x < 6) s[threadIdx.x] = threadIdx.x < 3 ? 0xffffffffu : 0u;
    __syncthreads();
    for (int a = 0; a < 3; ++a) {
        atomicMin(&s[a], f2o(mn[a]));
        atomicMax(&s[3 + a], f2o(mx[a]));
    }
    __syncthreads();
    if (threadIdx.x < 3) atomicMin(&acc[threadIdx.x], s[threadIdx.x]);
    else if (threadIdx.x < 6) atomicMax(&acc[threadIdx.x], s[threadIdx.x]);
}

__global__ void k_bbox_final(const uint32_t* acc, float* bbox)
{
    if (threadIdx.x < 6) bbox[threadIdx.x] = o2f(acc[threadIdx.x]);
}

// ---------------------------------------------------------------- encode
template <typename K>
struct MortonTraits;
template <>
struct MortonTraits<uint32_t> {
    static constexpr int kPasses = 4;
    __device__ static uint32_t code(float x, float y, float z)
    {
        // calculateMortonCode (Commons.cuh:497-510)
        uint32_t qx = (uint32_t)fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
        uint32_t qy = (uint32_t)fminf(fmaxf(y * 1024.0f, 0.0f), 1023.0f);
        uint32_t qz = (uint32_t)fminf(fmaxf(z * 1024.0f, 0.0f), 1023.0f);
        return spread10(qx) * 4u + spread10(qy) * 2u + spread10(qz);
    }
};
template <>
struct MortonTraits<uint64_t> {
    static constexpr int kPasses = 8;
    __device__ static uint64_t code(float x, float y, float z)
    {
        // calculateMortonCode64 (Commons.cuh:627-640)
        uint64_t qx = (uint64_t)fminf(fmaxf(x * 2097152.0f, 0.0f), 2097151.0f);
        uint64_t qy = (uint64_t)fminf(fmaxf(y * 2097152.0f, 0.0f), 2097151.0f);
        uint64_t qz = (uint64_t)fminf(fmaxf(z * 2097152.0f, 0.0f), 2097151.0f);
        return spread21(qx) * 4ull + spread21(qy) * 2ull + spread21(qz);
    }
};

// Leaf primitives of a build.  TriLeaves: the triangles of a scene, each
// leaf box the box of the three vertices (calculateLeafBoundingBox,
// LBVH.cu:430-446) and each leaf's data index the triangle.  BoxLeaves:
// caller-given primitive boxes, such as the fragments of split triangles,
// each naming the triangle it belongs to (the triangle-splitting variants
// generateMortonCodes*TSKernel LBVH.cu:99-161, buildTreeTSKernel :320-428
// and calculateNodeBoundingBoxesTSKernel :572-669): the leaf box is the
// given box (w dropped), the data index the given triangle.  Primitive p is
// the input position; the sort payload (the reference's routing index)
// carries it to the leaf.  Both give the key of the box centroid
// (max + min) * 0.5 (getTriangleCentroid :14-40, getBoundingBoxCentroid
// :42-51; x / 2 and x * 0.5 round identically).
struct TriLeaves {
    const float4* v;
    __device__ __forceinline__ void box(uint32_t p, float4& mn, float4& mx) const
    {
        const float4 a = v[3ull * p];
        const float4 b = v[3ull * p + 1];
        const float4 c = v[3ull * p + 2];
        mn = make_float4(fminf(fminf(a.x, b.x), c.x), fminf(fminf(a.y, b.y), c.y),
                         fminf(fminf(a.z, b.z), c.z), 0.0f);
        mx = make_float4(fmaxf(fmaxf(a.x, b.x), c.x), fmaxf(fmaxf(a.y, b.y), c.y),
                         fmaxf(fmaxf(a.z, b.z), c.z), 0.0f);
    }
    __device__ __forceinline__ int32_t data(uint32_t p) const { return (int32_t)p; }
};
struct BoxLeaves {
    const float4* bmin;
    const float4* bmax;
    const int32_t* tri;
    __device__ __forceinline__ void box(uint32_t p, float4& mn, float4& mx) const
    {
        const float4 a = bmin[p];
        const float4 b = bmax[p];
        mn = make_float4(a.x, a.y, a.z, 0.0f);
        mx = make_float4(b.x, b.y, b.z, 0.0f);
    }
    __device__ __forceinline__ int32_t data(uint32_t p) const { return tri[p]; }
};

// One thread per triangle, kEncItems triangles per thread.  Each wave keeps
// a private LDS histogram of every 8-bit digit of every sort pass; the block
// total is written to block_hist[block][pass][digit] (no global atomics).
// The same launch fills scratch the later stages need (slot words to ~0,
// counters to 0) so no separate memset is enqueued.
template <typename K, typename Src>
__global__ __launch_bounds__(kEncBlock) void k_encode(
    Src src, uint32_t n, const float* __restrict__ bbox, K* __restrict__ keys,
    uint32_t* __restrict__ idx, uint32_t* __restrict__ block_hist, uint64_t* __restrict__ fill64,
    uint32_t fill64_count, uint32_t* __restrict__ zero32, uint32_t zero32_count)
{
    constexpr int P = MortonTraits<K>::kPasses;
    __shared__ uint32_t hist[kEncBlock / 64][P * 256];
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    if (block_hist) {
        for (int k = tid; k < (kEncBlock / 64) * P * 256; k += kEncBlock) (&hist[0][0])[k] = 0;
    }
    __syncthreads();
    const float bminx = bbox[0], bminy = bbox[1], bminz = bbox[2];
    const float ex = bbox[3] - bminx, ey = bbox[4] - bminy, ez = bbox[5] - bminz;
    const uint32_t base = blockIdx.x * (uint32_t)kEncTile;
#pragma unroll 4
    for (int j = 0; j < kEncItems; ++j) {
        const uint32_t i = base + (uint32_t)j * kEncBlock + (uint32_t)tid;
        if (i < n) {
            // primitive box centroid (LBVH.cu:14-51), normalised (Commons.cuh:426-434)
            float4 bmn, bmx;
            src.box(i, bmn, bmx);
            const float cx = (bmx.x + bmn.x) * 0.5f;
            const float cy = (bmx.y + bmn.y) * 0.5f;
            const float cz = (bmx.z + bmn.z) * 0.5f;
            const K key = MortonTraits<K>::code((cx - bminx) / ex, (cy - bminy) / ey,
                                                (cz - bminz) / ez);
            keys[i] = key;
            if (idx) idx[i] = i;
            if (block_hist) {
#pragma unroll
                for (int p = 0; p < P; ++p)
                    atomicAdd(&hist[wave][p * 256 + (uint32_t)((key >> (8 * p)) & 255u)], 1u);
            }
        }
    }
    const uint32_t gsize = gridDim.x * kEncBlock;
    const uint32_t gtid = blockIdx.x * kEncBlock + tid;
    for (uint32_t k = gtid; k < fill64_count; k += gsize) fill64[k] = ~0ull;
    for (uint32_t k = gtid; k < zero32_count; k += gsize) zero32[k] = 0u;
    if (block_hist) {
        __syncthreads();
        uint32_t* out = block_hist + (size_t)blockIdx.x * P * 256;
        for (int e = tid; e < P * 256; e += kEncBlock) {
            uint32_t s = 0;
#pragma unroll
            for (int w = 0; w < kEncBlock / 64; ++w) s += hist[w][e];
            out[e] = s;
        }
    }
}

// ---------------------------------------------------------------- forests
// Batched builds keep many meshes in one mesh-major triangle array; seg[m] is
// the first triangle of mesh m and seg[M] the total.
__device__ __forceinline__ uint32_t mesh_of(const uint32_t* seg, uint32_t meshes, uint32_t i)
{
    uint32_t lo = 0, hi = meshes;  // seg[lo] <= i < seg[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (seg[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void k_fill_bbox_acc(uint32_t* acc, uint32_t meshes)
{
    for (uint32_t k = blockIdx.x * 256u + threadIdx.x; k < 6u * meshes; k += 256u * gridDim.x)
        acc[k] = (k % 6u) < 3u ? f2o(FLT_MAX) : f2o(FLT_MIN);
}

// Per-mesh scene boxes with SceneConverter's rule (SceneConverter.cpp:28-64).
// Each wave walks a contiguous run of triangles, each lane keeps a running box
// for its current mesh and flushes it with orderable-uint atomics when the
// mesh changes and at the end of the run (min/max are exact, so the order of
// the atomics does not matter).
constexpr int kBoxRun = 64 * 32;  // triangles per wave run
__global__ __launch_bounds__(256) void k_bbox_forest(const float4* __restrict__ v,
                                                     const uint32_t* __restrict__ seg,
                                                     uint32_t meshes, uint32_t n, uint32_t* acc)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t begin = wave * (uint32_t)kBoxRun;
    if (begin >= n) return;
    uint32_t m = mesh_of(seg, meshes, begin + lane < n ? begin + lane : begin);
    uint32_t next = seg[m + 1];
    float mn0 = FLT_MAX, mn1 = FLT_MAX, mn2 = FLT_MAX;
    float mx0 = FLT_MIN, mx1 = FLT_MIN, mx2 = FLT_MIN;
    bool any = false;
    for (uint32_t j = 0; j < (uint32_t)kBoxRun / 64u; ++j) {
        const uint32_t i = begin + j * 64u + lane;
        if (i >= n) break;
        if (i >= next) {
            if (any) {
                atomicMin(&acc[6 * m + 0], f2o(mn0));
                atomicMin(&acc[6 * m + 1], f2o(mn1));
                atomicMin(&acc[6 * m + 2], f2o(mn2));
                atomicMax(&acc[6 * m + 3], f2o(mx0));
                atomicMax(&acc[6 * m + 4], f2o(mx1));
                atomicMax(&acc[6 * m + 5], f2o(mx2));
            }
            while (i >= next) next = seg[++m + 1];
            mn0 = mn1 = mn2 = FLT_MAX;
            mx0 = mx1 = mx2 = FLT_MIN;
            any = false;
        }
        const float4 a = v[3ull * i];
        const float4 b = v[3ull * i + 1];
        const float4 c = v[3ull * i + 2];
        mn0 = fminf(mn0, fminf(fminf(a.x, b.x), c.x));
        mn1 = fminf(mn1, fminf(fminf(a.y, b.y), c.y));
        mn2 = fminf(mn2, fminf(fminf(a.z, b.z), c.z));
        mx0 = fmaxf(mx0, fmaxf(fmaxf(a.x, b.x), c.x));
        mx1 = fmaxf(mx1, fmaxf(fmaxf(a.y, b.y), c.y));
        mx2 = fmaxf(mx2, fmaxf(fmaxf(a.z, b.z), c.z));
        any = true;
    }
    // common case: the whole wave ends in one mesh -> one set of atomics
    const uint32_t m0 = __shfl(m, 0, 64);
    const bool uniform = __ballot(m != m0 && any) == 0ull;
    if (uniform) {
        for (int o = 32; o > 0; o >>= 1) {
            mn0 = fminf(mn0, __shfl_xor(mn0, o, 64));
            mn1 = fminf(mn1, __shfl_xor(mn1, o, 64));
            mn2 = fminf(mn2, __shfl_xor(mn2, o, 64));
            mx0 = fmaxf(mx0, __shfl_xor(mx0, o, 64));
            mx1 = fmaxf(mx1, __shfl_xor(mx1, o, 64));
            mx2 = fmaxf(mx2, __shfl_xor(mx2, o, 64));
        }
        any = lane == 0u;
        m = m0;
    }
    if (any) {
        atomicMin(&acc[6 * m + 0], f2o(mn0));
        atomicMin(&acc[6 * m + 1], f2o(mn1));
        atomicMin(&acc[6 * m + 2], f2o(mn2));
        atomicMax(&acc[6 * m + 3], f2o(mx0));
        atomicMax(&acc[6 * m + 4], f2o(mx1));
        atomicMax(&acc[6 * m + 5], f2o(mx2));
    }
}

__global__ void k_bbox_forest_final(const uint32_t* acc, uint32_t meshes, float* bbox)
{
    for (uint32_t k = blockIdx.x * 256u + threadIdx.x; k < 6u * meshes; k += 256u * gridDim.x)
        bbox[k] = o2f(acc[k]);
}

// Forest keys: mesh << 30 | 30-bit Morton code of the triangle in its own
// mesh's box, and the per-block digit histograms of `passes` 8-bit digits.
template <int ITEMS>
__global__ __launch_bounds__(kEncBlock) void k_encode_forest(
    const float4* __restrict__ v, const uint32_t* __restrict__ seg, uint32_t meshes, uint32_t n,
    const float* __restrict__ bbox, int passes, uint64_t* __restrict__ keys,
    uint32_t* __restrict__ block_hist, uint64_t* __restrict__ fill64, uint32_t fill64_count)
{
    __shared__ uint32_t hist[kEncBlock / 64][8 * 256];
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    for (int k = tid; k < (kEncBlock / 64) * passes * 256; k += kEncBlock) {
        const int w = k / (passes * 256);
        hist[w][k - w * passes * 256] = 0;
    }
    __syncthreads();
    const uint32_t base = blockIdx.x * (uint32_t)(kEncBlock * ITEMS);
    const uint32_t first = base + (uint32_t)tid;
    uint32_t m = first < n ? mesh_of(seg, meshes, first) : 0u;
    uint32_t next = seg[m + 1];
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t i = base + (uint32_t)j * kEncBlock + (uint32_t)tid;
        if (i >= n) break;
        while (i >= next) next = seg[++m + 1];
        const float* bb = bbox + 6ull * m;
        const float4 a = v[3ull * i];
        const float4 b = v[3ull * i + 1];
        const float4 c = v[3ull * i + 2];
        float cx = (fmaxf(fmaxf(a.x, b.x), c.x) + fminf(fminf(a.x, b.x), c.x)) * 0.5f;
        float cy = (fmaxf(fmaxf(a.y, b.y), c.y) + fminf(fminf(a.y, b.y), c.y)) * 0.5f;
        float cz = (fmaxf(fmaxf(a.z, b.z), c.z) + fminf(fminf(a.z, b.z), c.z)) * 0.5f;
        const uint32_t code = MortonTraits<uint32_t>::code(
            (cx - bb[0]) / (bb[3] - bb[0]), (cy - bb[1]) / (bb[4] - bb[1]),
            (cz - bb[2]) / (bb[5] - bb[2]));
        const uint64_t key = ((uint64_t)m << 30) | code;
        keys[i] = key;
        for (int p = 0; p < passes; ++p)
            atomicAdd(&hist[wave][p * 256 + (uint32_t)((key >> (8 * p)) & 255u)], 1u);
    }
    const uint32_t gsize = gridDim.x * kEncBlock;
    const uint32_t gtid = blockIdx.x * kEncBlock + tid;
    for (uint32_t k = gtid; k < fill64_count; k += gsize) fill64[k] = ~0ull;
    __syncthreads();
    uint32_t* out = block_hist + (size_t)blockIdx.x * passes * 256;
    for (int e = tid; e < passes * 256; e += kEncBlock) {
        uint32_t s = 0;
#pragma unroll
        for (int w = 0; w < kEncBlock / 64; ++w) s += hist[w][e];
        out[e] = s;
    }
}

// Global digit counts for every pass: workgroup (c, p) adds the histograms of
// encode blocks [c*kDigitChunk, (c+1)*kDigitChunk) for pass p, one thread per
// digit, with all of a thread's loads in flight at once, and adds its sum to
// totals[p][d] (zeroed beforehand; integer atomics, so the order is free).
constexpr int kDigitChunk = 16;
__global__ __launch_bounds__(256) void k_digit_totals(const uint32_t* __restrict__ block_hist,
                                                      uint32_t blocks, int passes,
                                                      uint32_t* __restrict__ totals)
{
    const int p = blockIdx.y;
    const int d = threadIdx.x;
    const uint32_t b0 = blockIdx.x * (uint32_t)kDigitChunk;
    uint32_t v[kDigitChunk];
#pragma unroll
    for (int k = 0; k < kDigitChunk; ++k) {
        const uint32_t b = b0 + (uint32_t)k;
        v[k] = b < blocks ? block_hist[((size_t)b * passes + p) * 256 + d] : 0u;
    }
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kDigitChunk; ++k) s += v[k];
    if (s) atomicAdd(&totals[p * 256 + d], s);
}

// In place: totals[p][d] -> number of keys whose digit p is smaller than d.
__global__ __launch_bounds__(256) void k_digit_scan(uint32_t* __restrict__ offsets)
{
    __shared__ uint32_t tmp[4];
    const int p = blockIdx.x;
    const int d = threadIdx.x;
    const uint32_t s = offsets[p * 256 + d];
    const uint32_t ex = block256_exclusive_sum(s, tmp, nullptr);
    offsets[p * 256 + d] = ex;
}

// Histogram of one digit position straight from keys (used when the keys
// did not come from the encode kernel).
template <typename K>
__global__ __launch_bounds__(256) void k_histogram(const K* __restrict__ keys, uint32_t n,
                                                   int passes, uint32_t* __restrict__ block_hist)
{
    __shared__ uint32_t hist[4][8 * 256];
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    for (int k = tid; k < 4 * 8 * 256; k += 256) (&hist[0][0])[k] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * (uint32_t)kEncTile;
    for (int j = 0; j < kEncItems; ++j) {
        const uint32_t i = base + (uint32_t)j * 256 + tid;
        if (i < n) {
            const K key = keys[i];
            for (int p = 0; p < passes; ++p)
                atomicAdd(&hist[wave][p * 256 + (uint32_t)((key >> (8 * p)) & 255u)], 1u);
        }
    }
    __syncthreads();
    uint32_t* out = block_hist + (size_t)blockIdx.x * passes * 256;
    for (int e = tid; e < passes * 256; e += 256)
        out[e] = hist[0][e] + hist[1][e] + hist[2][e] + hist[3][e];
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t count)
{
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < count; i += 256ull * gridDim.x)
        p[i] = v;
}

// Zeroes a[0, na) and b[0, nb) in one launch.
__global__ void k_zero_two(uint32_t* a, uint64_t na, uint32_t* b, uint64_t nb)
{
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < na + nb; i += 256ull * gridDim.x) {
        if (i < na) a[i] = 0u;
        else b[i - na] = 0u;
    }
}

// ---------------------------------------------------------------- onesweep
// One stable LSD pass over an 8-bit digit.  Tiles are taken in launch order
// from an atomic ticket; within a tile each wave ranks its ITEMS x 64 keys with
// ballot-based peer masks (keys keep their order: tile, wave, item, lane), a
// per-digit scan across waves gives tile-local positions (threads 0..255, one
// digit each), and a decoupled look-back over earlier tiles' published counts
// gives global positions.  Keys are reordered through LDS so writes of equal
// digits are contiguous.
template <typename K, int ITEMS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_onesweep(
    const K* __restrict__ kin, const uint32_t* __restrict__ vin, K* __restrict__ kout,
    uint32_t* __restrict__ vout, uint32_t n, int shift, const uint32_t* __restrict__ digit_off,
    uint32_t* status, uint32_t* tile_ticket, uint32_t* err, uint32_t spin_limit)
{
    static_assert(BLOCK % 256 == 0 && BLOCK * ITEMS == kSortTile, "tile shape");
    constexpr int WAVES = BLOCK / 64;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wave_hist[WAVES][256];
    __shared__ uint32_t s_local[256];
    __shared__ uint32_t s_global[256];
    __shared__ uint32_t s_tmp[4];
    __shared__ K s_keys[(BLOCK * ITEMS)];
    __shared__ uint32_t s_vals[(BLOCK * ITEMS)];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    if (tid == 0) s_tile = atomic_add_relaxed(tile_ticket, 1u);
    for (int k = tid; k < WAVES * 256; k += BLOCK) (&s_wave_hist[0][0])[k] = 0u;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t tile_base = tile * (uint32_t)(BLOCK * ITEMS);
    const uint32_t wbase = tile_base + (uint32_t)wave * (64u * ITEMS);

    K key[ITEMS];
    uint32_t val[ITEMS];
    uint32_t rank[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t e = wbase + (uint32_t)j * 64u + (uint32_t)lane;
        if (e < n) {
            key[j] = kin[e];
            val[j] = vin ? vin[e] : e;
        } else {
            key[j] = 0;
            val[j] = 0;
        }
    }
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t e = wbase + (uint32_t)j * 64u + (uint32_t)lane;
        const bool valid = e < n;
        const uint32_t d = (uint32_t)(key[j] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t r = 0;
        if (valid) {
            const uint32_t pre = (uint32_t)__popcll(peers & lt);
            const uint32_t before = s_wave_hist[wave][d];
            r = before + pre;
            if (pre == 0) s_wave_hist[wave][d] = before + (uint32_t)__popcll(peers);
        }
        rank[j] = r;
    }
    __syncthreads();

    // threads 0..255: thread == digit from here on
    const uint32_t d = (uint32_t)(tid & 255);
    const bool digit_thread = tid < 256;
    uint32_t count = 0;
    uint32_t* st = status + (size_t)tile * 256u;
    uint32_t inc = 0;
    if (digit_thread) {
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
            const uint32_t c = s_wave_hist[w][d];
            s_wave_hist[w][d] = count;
            count += c;
        }
        store_relaxed(&st[d], (tile == 0 ? kFlagInc : kFlagAgg) | count);
        inc = wave_inclusive_sum(count);
        if ((tid & 63) == 63) s_tmp[tid >> 6] = inc;
    }
    __syncthreads();
    if (digit_thread) {
        uint32_t add = 0;
        for (int k = 0; k < (tid >> 6); ++k) add += s_tmp[k];
        s_local[d] = add + inc - count;

        // Look back over the predecessors' status words for this digit, nearest
        // first, adding aggregates until an inclusive prefix: the words of up to
        // kLookWindow tiles are loaded together and consumed in order up to the
        // first one not yet published, which is then waited for.  Every
        // published word is final for the value it carries, so the sum is the
        // one-tile-at-a-time walk's.
        constexpr int kLookWindow = 8;
        uint32_t excl = 0;
        if (tile > 0) {
            int t = (int)tile - 1;
            uint32_t spins = 0;
            bool done = false;
            while (!done) {
                uint32_t w[kLookWindow];
#pragma unroll
                for (int k = 0; k < kLookWindow; ++k)
                    w[k] = t - k >= 0 ? load_relaxed(&status[(size_t)(t - k) * 256u + d]) : 0u;
                int used = 0;
                bool stall = false;
#pragma unroll
                for (int k = 0; k < kLookWindow; ++k) {
                    if (done || stall || t - k < 0) continue;
                    const uint32_t flag = w[k] & ~kValueMask;
                    if (flag == 0u) {
                        stall = true;
                        continue;
                    }
                    excl += w[k] & kValueMask;
                    ++used;
                    if (flag == kFlagInc) done = true;
                }
                t -= used;
                if (!done && stall) {
                    if (++spins > spin_limit) {
                        atomicOr(err, 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            store_relaxed(&st[d], kFlagInc | (excl + count));
        }
        s_global[d] = digit_off[d] + excl;
    }
    __syncthreads();

#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t e = wbase + (uint32_t)j * 64u + (uint32_t)lane;
        if (e < n) {
            const uint32_t dj = (uint32_t)(key[j] >> shift) & 255u;
            const uint32_t pos = s_local[dj] + s_wave_hist[wave][dj] + rank[j];
            s_keys[pos] = key[j];
            s_vals[pos] = val[j];
        }
    }
    __syncthreads();
    const uint32_t valid_count = min((uint32_t)(BLOCK * ITEMS), n - tile_base);
    for (uint32_t k = (uint32_t)tid; k < valid_count; k += BLOCK) {
        const K kk = s_keys[k];
        const uint32_t dd = (uint32_t)(kk >> shift) & 255u;
        const uint32_t o = s_global[dd] + (k - s_local[dd]);
        if (o < n) {  // always true unless a bounded look-back gave up (flagged in err)
            kout[o] = kk;
            vout[o] = s_vals[k];
        }
    }
}

// ---------------------------------------------------------------- build
// delta(i) = common-prefix length of sorted keys i and i+1, with equal keys
// falling back to the index prefix (longestCommonPrefix, LBVH.cu:163-201).
__device__ __forceinline__ int delta_at(const uint32_t* k, int i)
{
    const uint32_t a = k[i], b = k[i + 1];
    return a != b ? __clz(a ^ b) : 32 + __clz((uint32_t)(i ^ (i + 1)));
}
__device__ __forceinline__ int delta_at(const uint64_t* k, int i)
{
    const uint64_t a = k[i], b = k[i + 1];
    return a != b ? __clzll(a ^ b) : 64 + __clz((uint32_t)(i ^ (i + 1)));
}

// Fused radix-tree build and AABB refit.  One thread per leaf climbs: the node
// covering sorted range [lo, hi] has its parent split at hi (it is the left
// child) when delta(hi) > delta(lo-1), otherwise at lo-1 (right child).  In
// Karras numbering a left child is stored at its range's right end and a right
// child at its left end, the root at 0, so every internal node's index is
// known the moment it is completed.  The two children of a split meet through
// one 64-bit exchange (first arrival parks {far end, node}; the second takes
// it and builds the parent).
//
// A split whose node lies entirely inside this block's 256 leaves only ever
// sees this block's threads, so its exchange (and the box handed over) lives
// in LDS with workgroup-scope ordering.  Only splits whose node crosses a
// block boundary use the global slot, with the cross-XCD hand-over of
// device.hpp (coherent stores, relaxed agent-scope exchange, coherent loads;
// a thread peeks at the slot first and, when its sibling is already there,
// takes the parent at once).
// A split s is block-local iff the nearest smaller delta on both sides lies
// inside the block, i.e. the prefix minimum of delta over [base-1, s-1] and
// the suffix minimum over [s+1, base+255] are both below delta(s).
// Sorted-key views for the build.  delta(i) is the split strength between
// sorted positions i and i+1; -1 marks a wall that no node spans (the ends of
// the array, and in a forest the boundary between two meshes).
//
// Both views have a fallback form (`fb`), used when a bounded wait of the
// sort that produced the keys gave up (bit 0 of the context's status word):
// the sorted keys and indices may then be incomplete, so the build ignores
// them and makes the radix tree of the triangles in input order (every key
// taken as equal, so the split strengths come from positions alone).  The
// tree is well formed, so every later kernel stays in bounds, and the error
// is reported at the next synchronisation.
template <typename K>
struct SingleKeys {
    const K* k;
    int last;  // n - 1
    bool fb = false;
    __device__ __forceinline__ int delta(int i) const
    {
        if (i < 0 || i >= last) return -1;
        return fb ? 32 + __clz((uint32_t)(i ^ (i + 1))) : delta_at(k, i);
    }
};

// Forest of independent meshes sorted together: key = mesh << 30 | 30-bit
// Morton code, mesh-major.  Within a mesh delta equals the single-tree
// 30-bit delta with positions counted from the mesh's first position, so each
// mesh gets exactly the tree a build of that mesh alone gives.
struct ForestKeys {
    const uint64_t* k;
    int last;
    const uint32_t* seg_start;  // first sorted position of every mesh (meshes + 1 entries)
    uint32_t meshes;
    bool fb = false;
    __device__ __forceinline__ int delta(int i) const
    {
        if (i < 0 || i >= last) return -1;
        if (fb) {
            // mesh m holds positions seg_start[m] .. seg_start[m+1]-1
            uint32_t lo = 0, hi = meshes;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if ((int)seg_start[mid] <= i) lo = mid;
                else hi = mid;
            }
            if ((int)seg_start[lo + 1] == i + 1) return -1;
            const int s = (int)seg_start[lo];
            return 32 + __clz((uint32_t)((i - s) ^ (i + 1 - s)));
        }
        const uint64_t a = k[i], b = k[i + 1];
        if ((a ^ b) >> 30) return -1;
        const uint32_t x = (uint32_t)a ^ (uint32_t)b;
        if (x) return __clz(x);
        const int s = (int)seg_start[a >> 30];
        return 32 + __clz((uint32_t)((i - s) ^ (i + 1 - s)));
    }
};

// Node numbering: the node covering sorted range [lo, hi] is stored at hi when
// it is a left child, at lo when it is a right child, and a root (both ends at
// walls) at lo -- position 0 for a single tree, the mesh's first position in a
// forest.  Leaf of sorted position i is node first_leaf + i.
template <typename KV, typename Src>
__global__ __launch_bounds__(256) void k_build(KV kv_in, const uint32_t* __restrict__ idx,
                                               uint32_t n, int first_leaf,
                                               Src src, int32_t* parent,
                                               int32_t* left, int32_t* right, int32_t* data,
                                               float4* bbmin, float4* bbmax, float* area,
                                               int2* span, uint64_t* slots,
                                               const uint32_t* __restrict__ status)
{
    // bit 0 of the status word: a bounded wait of the sort gave up
    KV keys = kv_in;
    keys.fb = (status[0] & 1u) != 0u;
    __shared__ int s_d[257];
    __shared__ int s_wmin[8];
    __shared__ uint8_t s_local[256];
    __shared__ uint64_t s_slot[256];
    __shared__ float s_box[256][2][6];
    // Internal nodes made at block-local splits have indices inside this
    // block's range; their links and far range end are staged here and
    // written out, coalesced, when the block is done (scattered 4-byte stores
    // cost a whole memory sector each).
    __shared__ int32_t s_l[256];
    __shared__ int32_t s_r[256];
    __shared__ int32_t s_e[256];
    __shared__ uint8_t s_staged[256];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int base = (int)(blockIdx.x * 256u);
    // delta at positions base-1 .. base+255 (index t <-> position base-1+t)
    s_d[tid] = keys.delta(base - 1 + tid);
    if (tid == 0) s_d[256] = keys.delta(base + 255);
    s_slot[tid] = ~0ull;
    s_staged[tid] = 0;
    __syncthreads();
    // inclusive prefix minimum of s_d[0..tid] and suffix minimum of s_d[tid+1..256]
    int pm = s_d[tid];
    int sm = s_d[tid + 1];
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int a = __shfl_up(pm, o, 64);
        if (lane >= o) pm = min(pm, a);
        const int b = __shfl_down(sm, o, 64);
        if (lane + o < 64) sm = min(sm, b);
    }
    if (lane == 63) s_wmin[wave] = pm;
    if (lane == 0) s_wmin[4 + wave] = sm;
    __syncthreads();
    for (int w = 0; w < wave; ++w) pm = min(pm, s_wmin[w]);
    for (int w = wave + 1; w < 4; ++w) sm = min(sm, s_wmin[4 + w]);
    sm = min(sm, s_d[256]);
    __syncthreads();
    // split s = base + tid (tid < 255): left part s_d[0..tid] = positions base-1..s-1,
    // right part positions s+1..base+255 = s_d[tid+2..256]
    {
        const int ds = s_d[tid + 1];
        int right_min = __shfl_down(sm, 1, 64);
        // lane 63 needs the next wave's suffix: recompute from the wave minima
        if (lane == 63) {
            right_min = s_d[256];
            for (int w = wave + 1; w < 4; ++w) right_min = min(right_min, s_wmin[4 + w]);
        }
        s_local[tid] = (tid < 255 && pm < ds && right_min < ds) ? 1 : 0;
    }
    __syncthreads();

    const uint32_t i = (uint32_t)base + (uint32_t)tid;
    if (i < n) {
    // the primitive at sorted position i (its routing index) and its box
    const uint32_t prim = keys.fb ? i : idx[i];
    float4 mn, mx;
    src.box(prim, mn, mx);
    int node = first_leaf + (int)i;
    data[node] = src.data(prim);
    left[node] = -1;
    right[node] = -1;
    area[node] = box_area4(mn, mx);
    span[node] = make_int2((int)i, (int)i);

    // Every node's box is stored when its parent's split is known: with
    // device-coherent stores when that split crosses workgroups (the sibling
    // may read it from another XCD), with plain stores otherwise (the sibling
    // takes it from LDS; later kernels see it after this one ends).
    int lo = (int)i, hi = (int)i;
    int dl = keys.delta(lo - 1);
    int dr = keys.delta(hi);
    while (true) {
        if (dl < 0 && dr < 0) {
            parent[node] = -1;
            bbmin[node] = mn;
            bbmax[node] = mx;
            break;
        }
        const bool is_left = dr > dl;
        const int split = is_left ? hi : lo - 1;
        const uint32_t far_end = (uint32_t)(is_left ? lo : hi);
        const uint64_t mine = ((uint64_t)far_end << 32) | (uint32_t)node;
        const int k = split - base;
        uint64_t old;
        float4 smn, smx;
        if (k >= 0 && k < 255 && s_local[k]) {
            bbmin[node] = mn;
            bbmax[node] = mx;
            float* my = s_box[k][is_left ? 0 : 1];
            my[0] = mn.x; my[1] = mn.y; my[2] = mn.z;
            my[3] = mx.x; my[4] = mx.y; my[5] = mx.z;
            old = __hip_atomic_exchange(&s_slot[k], mine, __ATOMIC_ACQ_REL,
                                        __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old == ~0ull) break;
            const float* sb = s_box[k][is_left ? 1 : 0];
            smn = make_float4(sb[0], sb[1], sb[2], 0.0f);
            smx = make_float4(sb[3], sb[4], sb[5], 0.0f);
        } else {
            old = peek_crossing(&slots[split]);
            if (old == ~0ull) {
                st_coh(&bbmin[node], mn);
                st_coh(&bbmax[node], mx);
                old = exchange_crossing(&slots[split], mine);
                if (old == ~0ull) break;  // first arrival: the sibling finishes the parent
            } else {
                // the sibling is there: this thread makes the parent, so no
                // other thread reads this box during the launch
                bbmin[node] = mn;
                bbmax[node] = mx;
            }
            const int sib = (int)(uint32_t)old;
            smn = ld<true>(&bbmin[sib]);
            smx = ld<true>(&bbmax[sib]);
        }
        const int sib = (int)(uint32_t)old;
        const int sfar = (int)(uint32_t)(old >> 32);
        mn = make_float4(fminf(mn.x, smn.x), fminf(mn.y, smn.y), fminf(mn.z, smn.z), 0.0f);
        mx = make_float4(fmaxf(mx.x, smx.x), fmaxf(mx.y, smx.y), fmaxf(mx.z, smx.z), 0.0f);
        const int lchild = is_left ? node : sib;
        const int rchild = is_left ? sib : node;
        if (is_left) {
            hi = sfar;
            dr = keys.delta(hi);
        } else {
            lo = sfar;
            dl = keys.delta(lo - 1);
        }
        const int p = (dl < 0 && dr < 0) ? lo : (dr > dl ? hi : lo);
        if (k >= 0 && k < 255 && s_local[k]) {
            // p = lo or hi lies in this block's range; the rest at block end
            s_l[p - base] = lchild;
            s_r[p - base] = rchild;
            s_e[p - base] = p == lo ? hi : lo;
            s_staged[p - base] = 1;
        } else {
            left[p] = lchild;
            right[p] = rchild;
            data[p] = -1;
            parent[lchild] = p;
            parent[rchild] = p;
            area[p] = box_area4(mn, mx);
            span[p] = make_int2(lo, hi);
        }
        node = p;
    }
    }
    // Every staged node's box was stored by this block's thread that made it
    // (before that thread's next split), so after the barrier its area is
    // recomputed from the stored box, bit for bit the value in registers.
    __syncthreads();
    if (s_staged[tid]) {
        const int x = base + tid;
        const int l = s_l[tid], r = s_r[tid], e = s_e[tid];
        left[x] = l;
        right[x] = r;
        data[x] = -1;
        parent[l] = x;
        parent[r] = x;
        span[x] = make_int2(min(x, e), max(x, e));
        area[x] = box_area4(bbmin[x], bbmax[x]);
    }
}

// Bottom-up refit over a fixed topology (leaf boxes from the vertices, unions
// upwards), for trees whose vertices moved.
__global__ __launch_bounds__(256) void k_refit(uint32_t n, int first_leaf,
                                               const float4* __restrict__ v,
                                               const int32_t* __restrict__ parent,
                                               const int32_t* __restrict__ left,
                                               const int32_t* __restrict__ right,
                                               const int32_t* __restrict__ data, float4* bbmin,
                                               float4* bbmax, float* area, uint32_t* counters,
                                               const int2* __restrict__ span)
{
    __shared__ uint32_t s_cnt[256];
    s_cnt[threadIdx.x] = 0u;
    __syncthreads();
    const int blk = (int)blockIdx.x;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    int node = first_leaf + (int)i;
    const uint32_t tri = (uint32_t)data[node];
    const float4 a = v[3ull * tri];
    const float4 b = v[3ull * tri + 1];
    const float4 c = v[3ull * tri + 2];
    float4 mn = make_float4(fminf(fminf(a.x, b.x), c.x), fminf(fminf(a.y, b.y), c.y),
                            fminf(fminf(a.z, b.z), c.z), 0.0f);
    float4 mx = make_float4(fmaxf(fmaxf(a.x, b.x), c.x), fmaxf(fmaxf(a.y, b.y), c.y),
                            fmaxf(fmaxf(a.z, b.z), c.z), 0.0f);
    area[node] = box_area4(mn, mx);
    int cur = parent[node];
    // The next node's span, links and parent are loaded together as soon as
    // it is known (the topology is fixed).  A box is read during this launch
    // only by the sibling's thread at the parent; it needs a device-coherent
    // store only when that hand-over crosses workgroups.
    int2 nsp = make_int2(0, 0);
    int nl = -1, nr = -1, npp = -1;
    auto fetch = [&](int x) {
        nsp = span[x];
        nl = left[x];
        nr = right[x];
        npp = parent[x];
    };
    if (cur >= 0) fetch(cur);
    bool local = cur >= 0 && block_local(nsp, cur, blk);
    if (local || cur < 0) {
        bbmin[node] = mn;
        bbmax[node] = mx;
    } else {
        st_coh(&bbmin[node], mn);
        st_coh(&bbmax[node], mx);
    }
    while (cur >= 0) {
        uint32_t old;
        if (local) {
            old = arrive_local(&s_cnt[cur & 255]);
        } else {
            // a sibling already there (counter 1) is followed at once: its
            // box was stored before it arrived (see peek_crossing)
            old = ld<true>(&counters[cur]);
            if (old == 0u) old = arrive_crossing(&counters[cur]);
        }
        if (old == 0u) return;
        const int other = nl == node ? nr : nl;
        const float4 smn = local ? ld<false>(&bbmin[other]) : ld<true>(&bbmin[other]);
        const float4 smx = local ? ld<false>(&bbmax[other]) : ld<true>(&bbmax[other]);
        mn = make_float4(fminf(mn.x, smn.x), fminf(mn.y, smn.y), fminf(mn.z, smn.z), 0.0f);
        mx = make_float4(fmaxf(mx.x, smx.x), fmaxf(mx.y, smx.y), fmaxf(mx.z, smx.z), 0.0f);
        area[cur] = box_area4(mn, mx);
        node = cur;
        cur = npp;
        if (cur >= 0) fetch(cur);
        local = cur >= 0 && block_local(nsp, cur, blk);
        if (local || cur < 0) {
            bbmin[node] = mn;
            bbmax[node] = mx;
        } else {
            st_coh(&bbmin[node], mn);
            st_coh(&bbmax[node], mx);
        }
    }
}

}  // namespace

uint32_t encode_blocks(uint32_t n) { return div_up(n, kEncTile); }

atr_status run_scene_bbox(atr_context* ctx, const float4* verts, uint32_t n, float* d_bbox)
{
    uint32_t* acc = nullptr;
    ATR_TRY(dalloc_t(ctx, 8, &acc));
    ATR_TRY(launch(ctx, "bbox_init", 0, [&] { k_bbox_init<<<1, 64, 0, ctx->stream>>>(acc); }));
    const uint64_t count = 3ull * n;
    const uint32_t grid = std::min<uint32_t>(div_up(count, 256 * 8), 2048u);
    ATR_TRY(launch(ctx, "bbox", (double)count * 16.0, [&] {
        k_bbox<<<std::max(grid, 1u), 256, 0, ctx->stream>>>(verts, count, acc);
    }));
    ATR_TRY(launch(ctx, "bbox_final", 0, [&] { k_bbox_final<<<1, 64, 0, ctx->stream>>>(acc, d_bbox); }));
    dfree(ctx, acc);
    return ATR_OK;
}

atr_status run_encode(atr_context* ctx, const float4* verts, uint32_t n, const float* d_bbox,
                      int bits, void* keys, uint32_t* idx, uint32_t* block_hist,
                      uint32_t hist_blocks, uint64_t* fill64, uint32_t fill64_count,
                      uint32_t* zero32, uint32_t zero32_count)
{
    const uint32_t blocks = std::max(encode_blocks(n), 1u);
    if (block_hist && hist_blocks != blocks)
        return fail(ATR_ERR_INTERNAL, "encode histogram block count mismatch");
    const int kb = bits == 63 ? 8 : 4;
    const double bytes = (double)n * (48.0 + kb + (idx ? 4.0 : 0.0));
    const TriLeaves src{verts};
    if (bits == 63) {
        return launch(ctx, "encode63", bytes, [&] {
            k_encode<uint64_t><<<blocks, kEncBlock, 0, ctx->stream>>>(
                src, n, d_bbox, (uint64_t*)keys, idx, block_hist, fill64, fill64_count, zero32,
                zero32_count);
        });
    }
    return launch(ctx, "encode30", bytes, [&] {
        k_encode<uint32_t><<<blocks, kEncBlock, 0, ctx->stream>>>(
            src, n, d_bbox, (uint32_t*)keys, idx, block_hist, fill64, fill64_count, zero32,
            zero32_count);
    });
}

atr_status run_encode_boxes(atr_context* ctx, const float4* bmin, const float4* bmax, uint32_t n,
                            const float* d_bbox, int bits, void* keys, uint32_t* block_hist,
                            uint32_t hist_blocks, uint64_t* fill64, uint32_t fill64_count)
{
    const uint32_t blocks = std::max(encode_blocks(n), 1u);
    if (hist_blocks != blocks) return fail(ATR_ERR_INTERNAL, "encode histogram block count mismatch");
    const int kb = bits == 63 ? 8 : 4;
    const BoxLeaves src{bmin, bmax, nullptr};
    const double bytes = (double)n * (32.0 + kb);
    if (bits == 63) {
        return launch(ctx, "encode_boxes63", bytes, [&] {
            k_encode<uint64_t><<<blocks, kEncBlock, 0, ctx->stream>>>(
                src, n, d_bbox, (uint64_t*)keys, nullptr, block_hist, fill64, fill64_count,
                nullptr, 0);
        });
    }
    return launch(ctx, "encode_boxes30", bytes, [&] {
        k_encode<uint32_t><<<blocks, kEncBlock, 0, ctx->stream>>>(
            src, n, d_bbox, (uint32_t*)keys, nullptr, block_hist, fill64, fill64_count, nullptr,
            0);
    });
}

// Look-back status words of a sort (every tile shape holds kSortTile keys).
size_t sort_status_words(uint32_t n, int passes)
{
    return (size_t)passes * (std::max(div_up(n, (uint64_t)kSortTile), 1u) * 256u + 64u);
}

size_t sort_hist_words(uint32_t n, int passes)
{
    return (size_t)std::max(div_up(n, kEncTile), 1u) * passes * 256u;
}

// Stable LSD sort over `passes` 8-bit digits (the low 8*passes key bits).
// Results alternate between (keys, vals) and (keys_alt, vals_alt);
// *result_in_alt says where they end.  block_hist may come from the encode
// kernel; when NULL it is computed here.  work (optional) supplies the digit
// offsets (passes*256 words) and look-back status (sort_status_words) so that
// nothing is allocated (graph capture); otherwise they are allocated here.
atr_status run_sort(atr_context* ctx, int key_bits, int passes, void* keys, uint32_t* vals,
                    uint32_t n, void* keys_alt, uint32_t* vals_alt, bool vals_identity,
                    const uint32_t* block_hist, uint32_t hist_blocks, const SortWork* work,
                    bool* result_in_alt, int tile_items)
{
    const int kbytes = key_bits == 64 ? 8 : 4;
    if (passes < 1 || passes > kbytes) return fail(ATR_ERR_INTERNAL, "bad sort pass count");
    if (tile_items != 8 && tile_items != 16)
        return fail(ATR_ERR_INVALID_ARGUMENT, "sort tile items must be 8 or 16");
    const uint32_t tiles = std::max(div_up(n, (uint64_t)kSortTile), 1u);
    uint32_t* offsets = work ? work->offsets : nullptr;
    uint32_t* status = work ? work->status : nullptr;
    uint32_t* own_hist = nullptr;
    const size_t status_words = (size_t)passes * (tiles * 256u + 64u);
    if (!work) {
        ATR_TRY(dalloc_t(ctx, (size_t)passes * 256, &offsets));
        ATR_TRY(dalloc_t(ctx, status_words, &status));
    }
    uint32_t* hist_buf = nullptr;
    if (!block_hist) {
        hist_blocks = std::max(div_up(n, kEncTile), 1u);
        if (work && work->hist) {
            hist_buf = work->hist;
        } else {
            ATR_TRY(dalloc_t(ctx, (size_t)hist_blocks * passes * 256, &own_hist));
            hist_buf = own_hist;
        }
        if (key_bits == 64) {
            ATR_TRY(launch(ctx, "histogram64", (double)n * kbytes, [&] {
                k_histogram<uint64_t><<<hist_blocks, 256, 0, ctx->stream>>>(
                    (const uint64_t*)keys, n, passes, hist_buf);
            }));
        } else {
            ATR_TRY(launch(ctx, "histogram32", (double)n * kbytes, [&] {
                k_histogram<uint32_t><<<hist_blocks, 256, 0, ctx->stream>>>(
                    (const uint32_t*)keys, n, passes, hist_buf);
            }));
        }
        block_hist = hist_buf;
    }
    const uint64_t offset_words = (uint64_t)passes * 256;
    ATR_TRY(launch(ctx, "fill_status", (double)(status_words + offset_words) * 4.0, [&] {
        k_zero_two<<<std::min<uint32_t>(div_up(status_words + offset_words, 256), 1024u), 256, 0,
                     ctx->stream>>>(status, status_words, offsets, offset_words);
    }));
    ATR_TRY(launch(ctx, "digit_totals", (double)hist_blocks * passes * 1024.0, [&] {
        k_digit_totals<<<dim3(div_up(hist_blocks, kDigitChunk), passes), 256, 0, ctx->stream>>>(
            block_hist, hist_blocks, passes, offsets);
    }));
    ATR_TRY(launch(ctx, "digit_scan", (double)offset_words * 8.0, [&] {
        k_digit_scan<<<passes, 256, 0, ctx->stream>>>(offsets);
    }));
    void* kbuf[2] = {keys, keys_alt};
    uint32_t* vbuf[2] = {vals, vals_alt};
    int cur = 0;
    for (int p = 0; p < passes; ++p) {
        uint32_t* st = status + (size_t)p * (tiles * 256u + 64u);
        uint32_t* ticket = st + tiles * 256u;
        const uint32_t* vin = (p == 0 && vals_identity) ? nullptr : vbuf[cur];
        const double bytes = (double)n * 2.0 * (kbytes + 4.0) - (vin ? 0.0 : 4.0 * n);
        atr_status s;
        if (key_bits == 64) {
            s = launch(ctx, "onesweep64", bytes, [&] {
                if (tile_items == 16)
                    k_onesweep<uint64_t, 16, 512><<<tiles, 512, 0, ctx->stream>>>(
                        (const uint64_t*)kbuf[cur], vin, (uint64_t*)kbuf[1 - cur], vbuf[1 - cur],
                        n, 8 * p, offsets + p * 256, st, ticket, ctx->d_status, ctx->spin_limit);
                else
                    k_onesweep<uint64_t, 8, 1024><<<tiles, 1024, 0, ctx->stream>>>(
                        (const uint64_t*)kbuf[cur], vin, (uint64_t*)kbuf[1 - cur], vbuf[1 - cur],
                        n, 8 * p, offsets + p * 256, st, ticket, ctx->d_status, ctx->spin_limit);
            });
        } else {
            s = launch(ctx, "onesweep32", bytes, [&] {
                if (tile_items == 16)
                    k_onesweep<uint32_t, 16, 512><<<tiles, 512, 0, ctx->stream>>>(
                        (const uint32_t*)kbuf[cur], vin, (uint32_t*)kbuf[1 - cur], vbuf[1 - cur],
                        n, 8 * p, offsets + p * 256, st, ticket, ctx->d_status, ctx->spin_limit);
                else
                    k_onesweep<uint32_t, 8, 1024><<<tiles, 1024, 0, ctx->stream>>>(
                        (const uint32_t*)kbuf[cur], vin, (uint32_t*)kbuf[1 - cur], vbuf[1 - cur],
                        n, 8 * p, offsets + p * 256, st, ticket, ctx->d_status, ctx->spin_limit);
            });
        }
        ATR_TRY(s);
        cur = 1 - cur;
    }
    const bool in_alt = cur == 1;
    if (!work) {
        dfree(ctx, offsets);
        dfree(ctx, status);
    }
    if (own_hist) dfree(ctx, own_hist);
    *result_in_alt = in_alt;
    return ATR_OK;
}

// Forest stages (batched builds).
atr_status run_bbox_forest(atr_context* ctx, const float4* verts, const uint32_t* d_seg,
                           uint32_t meshes, uint32_t n, uint32_t* acc, float* d_bbox)
{
    const uint32_t g6 = std::max(std::min<uint32_t>(div_up(6ull * meshes, 256), 1024u), 1u);
    ATR_TRY(launch(ctx, "bbox_forest_init", meshes * 24.0, [&] {
        k_fill_bbox_acc<<<g6, 256, 0, ctx->stream>>>(acc, meshes);
    }));
    const uint32_t waves = div_up(n, kBoxRun);
    ATR_TRY(launch(ctx, "bbox_forest", (double)n * 48.0, [&] {
        k_bbox_forest<<<std::max(div_up(waves, 4), 1u), 256, 0, ctx->stream>>>(verts, d_seg,
                                                                                 meshes, n, acc);
    }));
    return launch(ctx, "bbox_forest_final", meshes * 48.0, [&] {
        k_bbox_forest_final<<<g6, 256, 0, ctx->stream>>>(acc, meshes, d_bbox);
    });
}

int forest_passes(uint32_t meshes)
{
    int mbits = 0;
    while (mbits < 34 && (1ull << mbits) < meshes) ++mbits;
    return (30 + mbits + 7) / 8;
}

uint32_t encode_forest_blocks(uint32_t n, int items)
{
    return std::max(div_up(n, (uint64_t)kEncBlock * items), 1u);
}

atr_status run_encode_forest(atr_context* ctx, const float4* verts, const uint32_t* d_seg,
                             uint32_t meshes, uint32_t n, const float* d_bbox, uint64_t* keys,
                             uint32_t* block_hist, uint64_t* fill64, uint32_t fill64_count,
                             int items)
{
    if (items != 8 && items != 16)
        return fail(ATR_ERR_INVALID_ARGUMENT, "encode items must be 8 or 16");
    const int passes = forest_passes(meshes);
    const uint32_t blocks = encode_forest_blocks(n, items);
    return launch(ctx, "encode_forest", (double)n * 56.0, [&] {
        if (items == 16)
            k_encode_forest<16><<<blocks, kEncBlock, 0, ctx->stream>>>(
                verts, d_seg, meshes, n, d_bbox, passes, keys, block_hist, fill64, fill64_count);
        else
            k_encode_forest<8><<<blocks, kEncBlock, 0, ctx->stream>>>(
                verts, d_seg, meshes, n, d_bbox, passes, keys, block_hist, fill64, fill64_count);
    });
}

atr_status run_build(atr_context* ctx, int bits, const void* keys, const uint32_t* idx,
                     uint32_t n, const float4* verts, atr_tree* t, uint64_t* slots)
{
    const uint32_t grid = div_up(n, 256);
    const double bytes = (double)n * 176.0;
    const int fl = (int)t->first_leaf;
    const TriLeaves src{verts};
    if (bits == 63) {
        SingleKeys<uint64_t> kv{(const uint64_t*)keys, (int)n - 1};
        return launch(ctx, "build_refit63", bytes, [&] {
            k_build<<<grid, 256, 0, ctx->stream>>>(kv, idx, n, fl, src, t->parent, t->left,
                                                   t->right, t->data, t->bbmin, t->bbmax, t->area,
                                                   t->span, slots, ctx->d_status);
        });
    }
    SingleKeys<uint32_t> kv{(const uint32_t*)keys, (int)n - 1};
    return launch(ctx, "build_refit30", bytes, [&] {
        k_build<<<grid, 256, 0, ctx->stream>>>(kv, idx, n, fl, src, t->parent, t->left,
                                               t->right, t->data, t->bbmin, t->bbmax, t->area,
                                               t->span, slots, ctx->d_status);
    });
}

atr_status run_build_boxes(atr_context* ctx, int bits, const void* keys, const uint32_t* idx,
                           uint32_t n, const float4* bmin, const float4* bmax,
                           const int32_t* tri, atr_tree* t, uint64_t* slots)
{
    const uint32_t grid = div_up(n, 256);
    // leaf: 4 routing + 32 box + 4 triangle id read; the rest as a triangle build
    const double bytes = (double)n * 164.0;
    const int fl = (int)t->first_leaf;
    const BoxLeaves src{bmin, bmax, tri};
    if (bits == 63) {
        SingleKeys<uint64_t> kv{(const uint64_t*)keys, (int)n - 1};
        return launch(ctx, "build_boxes63", bytes, [&] {
            k_build<<<grid, 256, 0, ctx->stream>>>(kv, idx, n, fl, src, t->parent, t->left,
                                                   t->right, t->data, t->bbmin, t->bbmax, t->area,
                                                   t->span, slots, ctx->d_status);
        });
    }
    SingleKeys<uint32_t> kv{(const uint32_t*)keys, (int)n - 1};
    return launch(ctx, "build_boxes30", bytes, [&] {
        k_build<<<grid, 256, 0, ctx->stream>>>(kv, idx, n, fl, src, t->parent, t->left,
                                               t->right, t->data, t->bbmin, t->bbmax, t->area,
                                               t->span, slots, ctx->d_status);
    });
}

atr_status run_build_forest(atr_context* ctx, const uint64_t* keys, const uint32_t* idx,
                            const float4* verts, atr_tree* t, const uint32_t* d_seg,
                            uint32_t meshes, uint64_t* slots)
{
    const uint32_t n = t->n;
    ForestKeys kv{keys, (int)n - 1, d_seg, meshes};
    return launch(ctx, "build_refit_forest", (double)n * 176.0, [&] {
        k_build<<<div_up(n, 256), 256, 0, ctx->stream>>>(
            kv, idx, n, (int)t->first_leaf, TriLeaves{verts}, t->parent, t->left, t->right, t->data,
            t->bbmin, t->bbmax, t->area, t->span, slots, ctx->d_status);
    });
}

atr_status run_refit(atr_context* ctx, atr_tree* t, const float4* verts, uint32_t* counters)
{
    const uint32_t n = t->n;
    ATR_TRY(launch(ctx, "fill_counters", (double)t->first_leaf * 4.0, [&] {
        k_fill_u32<<<std::max(std::min<uint32_t>(div_up(t->first_leaf, 256), 1024u), 1u), 256, 0,
                     ctx->stream>>>(counters, 0u, (uint64_t)t->first_leaf);
    }));
    return launch(ctx, "refit", (double)n * 176.0, [&] {
        k_refit<<<div_up(n, 256), 256, 0, ctx->stream>>>(n, (int)t->first_leaf, verts, t->parent,
                                                         t->left, t->right, t->data, t->bbmin,
                                                         t->bbmax, t->area, counters, t->span);
    });
}

}  // namespace atr
