// Device helpers for the gfx950 kernels.  All translation units are compiled
// with -ffp-contract=off so every SAH / area expression below is evaluated as
// separate IEEE fp32 multiplies and adds, identical to the host oracle.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace atr {

constexpr int kWave = 64;

// 2*(x*y + x*z + y*z) of a box (BVHRT-Core/src/Commons.cuh:99-125).
__device__ __forceinline__ float box_area(float mnx, float mny, float mnz, float mxx, float mxy,
                                          float mxz)
{
    float sx = mxx - mnx;
    float sy = mxy - mny;
    float sz = mxz - mnz;
    return 2.0f * (sx * sy + sx * sz + sy * sz);
}

__device__ __forceinline__ float box_area4(const float4& mn, const float4& mx)
{
    return box_area(mn.x, mn.y, mn.z, mx.x, mx.y, mx.z);
}

// Area of the union of two boxes (Commons.cuh:156-164).
__device__ __forceinline__ float union_area(float amnx, float amny, float amnz, float amxx,
                                            float amxy, float amxz, float bmnx, float bmny,
                                            float bmnz, float bmxx, float bmxy, float bmxz)
{
    float sx = fmaxf(amxx, bmxx) - fminf(amnx, bmnx);
    float sy = fmaxf(amxy, bmxy) - fminf(amny, bmny);
    float sz = fmaxf(amxz, bmxz) - fminf(amnz, bmnz);
    return 2.0f * (sx * sy + sx * sz + sy * sz);
}

// 10-bit and 21-bit bit spreading for Morton interleave (Commons.cuh:481-488,
// 606-618): every input bit moves to position 3*b.
__device__ __forceinline__ uint32_t spread10(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__device__ __forceinline__ uint64_t spread21(uint64_t v)
{
    v &= 0x1fffffull;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

// --------------------------------------------------------------- atomics
// Agent-scope (whole GPU, all XCDs) relaxed operations.  (An acq_rel RMW at
// this scope would bracket the atomic with an L2 writeback and invalidate on
// gfx950; the bottom-up protocols below avoid that, see the hand-over notes.)
__device__ __forceinline__ uint32_t load_relaxed(const uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_relaxed(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t atomic_add_relaxed(uint32_t* p, uint32_t v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------- cross-XCD hand-over
// Bottom-up passes hand results from one climbing thread to another.  When
// both belong to the same workgroup (a "block-local" node) a workgroup-scope
// LDS atomic orders them.  When the node's subtree spans several workgroups
// the two threads may sit on different XCDs, whose L2 caches are not kept
// coherent; an agent-scope acquire/release there would write back and
// invalidate the whole XCD L2 on every hand-over.  Instead:
//   - values a thread on another XCD may read later are written with
//     device-coherent stores (agent-scope relaxed atomic stores, which go
//     through to the point of coherence shared by all XCDs);
//   - before every arrival a thread waits until all of its wave's memory
//     operations have completed (so those stores are visible device-wide);
//   - the crossing arrival itself is a relaxed agent-scope atomic, and the
//     second arrival reads the handed-over values with device-coherent loads
//     issued only after the atomic has returned.
__device__ __forceinline__ void wait_memory()
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ void st_coh(int32_t* p, int32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coh(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coh(float* p, float v)
{
    __hip_atomic_store((uint32_t*)p, __float_as_uint(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coh(int2* p, int2 v)
{
    const uint64_t w = ((uint64_t)(uint32_t)v.y << 32) | (uint32_t)v.x;
    __hip_atomic_store((uint64_t*)p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coh(float4* p, float4 v)
{
    uint64_t* q = (uint64_t*)p;
    const uint64_t a = ((uint64_t)__float_as_uint(v.y) << 32) | __float_as_uint(v.x);
    const uint64_t b = ((uint64_t)__float_as_uint(v.w) << 32) | __float_as_uint(v.z);
    __hip_atomic_store(q, a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 1, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_coh(int4* p, int4 v)
{
    uint64_t* q = (uint64_t*)p;
    const uint64_t a = ((uint64_t)(uint32_t)v.y << 32) | (uint32_t)v.x;
    const uint64_t b = ((uint64_t)(uint32_t)v.w << 32) | (uint32_t)v.z;
    __hip_atomic_store(q, a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 1, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Loads: COH selects the device-coherent flavour (crossing readers).
template <bool COH>
__device__ __forceinline__ int32_t ld(const int32_t* p)
{
    if constexpr (COH)
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        return *p;
}
template <bool COH>
__device__ __forceinline__ uint32_t ld(const uint32_t* p)
{
    if constexpr (COH)
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        return *p;
}
template <bool COH>
__device__ __forceinline__ float ld(const float* p)
{
    if constexpr (COH)
        return __uint_as_float(
            __hip_atomic_load((const uint32_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    else
        return *p;
}
template <bool COH>
__device__ __forceinline__ int2 ld(const int2* p)
{
    if constexpr (COH) {
        const uint64_t w =
            __hip_atomic_load((const uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return make_int2((int)(uint32_t)w, (int)(uint32_t)(w >> 32));
    } else {
        return *p;
    }
}
template <bool COH>
__device__ __forceinline__ float4 ld(const float4* p)
{
    if constexpr (COH) {
        const uint64_t* q = (const uint64_t*)p;
        const uint64_t a = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t b = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return make_float4(__uint_as_float((uint32_t)a), __uint_as_float((uint32_t)(a >> 32)),
                           __uint_as_float((uint32_t)b), __uint_as_float((uint32_t)(b >> 32)));
    } else {
        return *p;
    }
}

template <bool COH>
__device__ __forceinline__ int4 ld(const int4* p)
{
    if constexpr (COH) {
        const uint64_t* q = (const uint64_t*)p;
        const uint64_t a = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t b = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return make_int4((int)(uint32_t)a, (int)(uint32_t)(a >> 32), (int)(uint32_t)b,
                         (int)(uint32_t)(b >> 32));
    } else {
        return *p;
    }
}

// Arrival at a node whose two children were produced inside this workgroup.
__device__ __forceinline__ uint32_t arrive_local(uint32_t* lds_counter)
{
    wait_memory();
    return __hip_atomic_fetch_add(lds_counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Arrival at a node whose children may have been produced on other XCDs.
__device__ __forceinline__ uint32_t arrive_crossing(uint32_t* counter)
{
    wait_memory();
    const uint32_t old = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
    wait_memory();
    return old;
}

__device__ __forceinline__ uint64_t exchange_crossing(uint64_t* slot, uint64_t v)
{
    wait_memory();
    const uint64_t old = __hip_atomic_exchange(slot, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    wait_memory();
    return old;
}

// Peek before arriving at a crossing slot.  A thread whose sibling has
// already arrived sees the sibling's value with one device-coherent load and
// continues without writing the slot or first waiting for its own stores: no
// third thread arrives there, and the sibling published only after its own
// stores had completed, so loads issued once the peeked value has returned
// (everything that depends on it) see them.  This takes the store wait and
// the atomic off the path of the continuing thread, which is the later of
// the two and hence the one the chain to the root waits for.  Returns the
// empty value (all ones) when the sibling has not arrived; the thread then
// arrives with exchange_crossing as before.
__device__ __forceinline__ uint64_t peek_crossing(const uint64_t* slot)
{
    return __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Block-local test: every leaf under the node (span) and the node index itself
// lie in this workgroup's block of 2^SHIFT leaves (256 unless stated).
template <int SHIFT = 8>
__device__ __forceinline__ bool block_local(int2 span, int node, int blk)
{
    return (span.x >> SHIFT) == blk && (span.y >> SHIFT) == blk && (node >> SHIFT) == blk;
}

// --------------------------------------------------------------- wave ops
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt()
{
    return (1ull << (uint32_t)__lane_id()) - 1ull;
}

// Inclusive prefix sum across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v)
{
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive prefix sum over a 256-thread block (one value per thread).
// `tmp` must hold 4 words of LDS.  Contains __syncthreads().
__device__ __forceinline__ uint32_t block256_exclusive_sum(uint32_t v, uint32_t* tmp,
                                                           uint32_t* total)
{
    const int tid = threadIdx.x;
    uint32_t inc = wave_inclusive_sum(v);
    if ((tid & 63) == 63) tmp[tid >> 6] = inc;
    __syncthreads();
    uint32_t add = 0;
    const int w = tid >> 6;
    for (int k = 0; k < w; ++k) add += tmp[k];
    if (total) *total = tmp[0] + tmp[1] + tmp[2] + tmp[3];
    __syncthreads();
    return add + inc - v;
}

// --------------------------------------------------------- ray distribution
// Rays for persistent traversal waves.  kRayCounters counters, each on its own
// 256-byte line, hand out 64-ray blocks interleaved over the ray range
// (counter k serves blocks k, k + 8, k + 16, ...), a chunk of positions per
// atomic into a wave-local queue from which the wave's idle lanes take rays.
// A single counter taking at most 64 rays per atomic serialised the launch
// (4,194,304 rays that end at once took 0.85 ms on one MI355X, 0.18 ms this
// way); spreading the atomics over eight addresses and taking up to two blocks
// at a time removes that, and the interleave keeps every counter's share of
// the image alike.  The chunk shrinks with the rays per wave (about four
// chunks per wave, 16 to 128 positions) so that few rays still spread over
// every wave.  A wave
// starts on counter blockIdx.x % 8 and moves on when its counter runs dry, so
// every ray is handed out exactly once and the wave stops only after it has
// seen all eight counters dry.  Counters start at zero.
constexpr int kRayCounters = 8;
constexpr uint32_t kRayCounterStride = 64;  // words between counters
constexpr uint32_t kRayCounterWords = kRayCounters * kRayCounterStride;

struct RayQueue {
    uint32_t* counters;
    uint32_t R;
    int ck;
    int dry = 0;
    uint32_t next = 0, end = 0;
    uint32_t chunk;
    bool exhausted = false;

    __device__ __forceinline__ RayQueue(uint32_t* c, uint32_t r)
        : counters(c), R(r), ck((int)(blockIdx.x % kRayCounters))
    {
        const uint32_t waves = gridDim.x * (blockDim.x / 64u);
        chunk = min(128u, max(16u, r / (4u * waves)));
    }
    // positions of counter k's sequence: 64 per block it serves
    __device__ __forceinline__ uint32_t span(int k) const
    {
        const uint32_t nb = (R + 63u) / 64u;
        return nb > (uint32_t)k ? ((nb - 1u - (uint32_t)k) / kRayCounters + 1u) * 64u : 0u;
    }
    // ray index of position s of the current counter's sequence (may be >= R
    // in the last, partial block: such positions carry no ray)
    __device__ __forceinline__ uint32_t ray_of(uint32_t s) const
    {
        return ((s >> 6) * kRayCounters + (uint32_t)ck) * 64u + (s & 63u);
    }
    // Wave-uniform: up to `want` positions [first, first + return value) of
    // the current counter; 0 while moving to another counter.
    __device__ __forceinline__ uint32_t take(uint32_t want, int leader, uint32_t& first)
    {
        if (next == end) {
            const uint32_t sk = span(ck);
            uint32_t base = 0;
            if ((int)__lane_id() == leader)
                base = atomicAdd(&counters[(uint32_t)ck * kRayCounterStride], chunk);
            base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
            if (base >= sk) {
                if (++dry == kRayCounters) exhausted = true;
                else ck = (ck + 1) % kRayCounters;
                return 0u;
            }
            next = base;
            end = min(base + chunk, sk);
        }
        const uint32_t t = min(want, end - next);
        first = next;
        next += t;
        return t;
    }
};

// --------------------------------------------------------- traversal stack
// Per-lane stack of a traversal: the newest RING entries live in an LDS ring
// laid out [entry][thread] (a wave's lanes touch consecutive words), older
// entries spill to a private array of SPILL entries, which only deep
// traversals touch.  `ring` points at this thread's column of the block's
// RING * BLOCK words.  push() refuses (returns false) beyond RING + SPILL
// entries instead of writing out of bounds.
template <int RING, int SPILL, int BLOCK>
struct LdsStack {
    uint32_t* ring;
    uint32_t spill[SPILL > 0 ? SPILL : 1];
    int n = 0;

    __device__ __forceinline__ explicit LdsStack(uint32_t* column) : ring(column) {}
    __device__ __forceinline__ bool push(uint32_t v)
    {
        if (n >= RING) {
            const int old = n - RING;
            if (old >= SPILL) return false;
            spill[old] = ring[(old % RING) * BLOCK];
        }
        ring[(n % RING) * BLOCK] = v;
        ++n;
        return true;
    }
    __device__ __forceinline__ uint32_t pop()
    {
        --n;
        const uint32_t v = ring[(n % RING) * BLOCK];
        if (n >= RING) {
            const int old = n - RING;
            ring[(old % RING) * BLOCK] = spill[old];
        }
        return v;
    }
};

}  // namespace atr
