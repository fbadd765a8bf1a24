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
// Agent-scope (whole GPU, all XCDs) orderings.  On gfx950 an acq_rel RMW
// brackets the atomic with an L2 writeback (release) and invalidate
// (acquire), which the cross-XCD bottom-up protocols need.
__device__ __forceinline__ uint32_t atomic_add_acq_rel(uint32_t* p, uint32_t v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t atomic_xchg_acq_rel(uint64_t* p, uint64_t v)
{
    return __hip_atomic_exchange(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}

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

// Maximum spin iterations of a bounded wait before it gives up and flags an
// error (a correct launch never gets close).
constexpr uint32_t kSpinLimit = 1u << 26;

}  // namespace atr
