// Ray coherence sort on gfx950: 6D Morton keys of rays (origin in the rays'
// bounding box, direction on the unit cube) and a stable sort of the rays by
// them, after RayBuffer::mortonSort (gpu-ray-traversal/src/rt/ray/
// RayBuffer.cpp:103-163) and its kernels findAABBKernel, genMortonKeysKernel
// and reorderRaysKernel (rt/ray/RayBufferKernels.cu:64-196).  The reference
// builds the keys on the GPU and sorts them on the CPU; here the 192-bit keys
// are held as three 64-bit words and sorted on the GPU by three stable LSD
// radix sorts (low word first) with the build's onesweep sort.
#include "device.hpp"
#include "internal.hpp"

namespace atr {

namespace {

// findAABBKernel (RayBufferKernels.cu:64-135): the box of every ray's origin
// and end point o + d * tmax.  An end point that is not finite (tmax = inf,
// the convention of closest-hit rays without a far plane) is left out; the
// reference's rays always carry a finite tmax.  Each block reduces its rays
// (waves by cross-lane min / max, then the block's waves through LDS) to one
// partial box; a single block then reduces the partials.  min / max are
// exact, so the order of the reduction does not matter.
constexpr int kBoxBlocks = 1024;
__global__ __launch_bounds__(256) void k_ray_box(const float4* __restrict__ rays, uint32_t r,
                                                 float* __restrict__ partial)
{
    __shared__ float s[4][6];
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < r; i += 256u * gridDim.x) {
        const float4 o = rays[2ull * i];
        const float4 d = rays[2ull * i + 1];
        mn[0] = fminf(mn[0], o.x);
        mn[1] = fminf(mn[1], o.y);
        mn[2] = fminf(mn[2], o.z);
        mx[0] = fmaxf(mx[0], o.x);
        mx[1] = fmaxf(mx[1], o.y);
        mx[2] = fmaxf(mx[2], o.z);
        const float ex = o.x + d.x * d.w;
        const float ey = o.y + d.y * d.w;
        const float ez = o.z + d.z * d.w;
        if (isfinite(ex) && isfinite(ey) && isfinite(ez)) {
            mn[0] = fminf(mn[0], ex);
            mn[1] = fminf(mn[1], ey);
            mn[2] = fminf(mn[2], ez);
            mx[0] = fmaxf(mx[0], ex);
            mx[1] = fmaxf(mx[1], ey);
            mx[2] = fmaxf(mx[2], ez);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        for (int a = 0; a < 3; ++a) {
            mn[a] = fminf(mn[a], __shfl_xor(mn[a], o, 64));
            mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], o, 64));
        }
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0u) {
        for (int a = 0; a < 3; ++a) {
            s[wave][a] = mn[a];
            s[wave][3 + a] = mx[a];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        float v = s[0][a];
        for (int w = 1; w < 4; ++w) v = a < 3 ? fminf(v, s[w][a]) : fmaxf(v, s[w][a]);
        partial[6ull * blockIdx.x + a] = v;
    }
}

// The partial boxes of `blocks` blocks to the rays' box (one block).
__global__ __launch_bounds__(256) void k_ray_box_final(const float* __restrict__ partial,
                                                       uint32_t blocks, float* __restrict__ box)
{
    __shared__ float s[4][6];
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t b = threadIdx.x; b < blocks; b += 256u) {
        for (int a = 0; a < 3; ++a) {
            mn[a] = fminf(mn[a], partial[6ull * b + a]);
            mx[a] = fmaxf(mx[a], partial[6ull * b + 3 + a]);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        for (int a = 0; a < 3; ++a) {
            mn[a] = fminf(mn[a], __shfl_xor(mn[a], o, 64));
            mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], o, 64));
        }
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0u) {
        for (int a = 0; a < 3; ++a) {
            s[wave][a] = mn[a];
            s[wave][3 + a] = mx[a];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        float v = s[0][a];
        for (int w = 1; w < 4; ++w) v = a < 3 ? fminf(v, s[w][a]) : fmaxf(v, s[w][a]);
        box[a] = v;
    }
}

// float -> uint32 with the saturation of the GPU conversion the reference
// relies on: NaN and values below 1 give 0, values of 2^32 and above give
// 0xFFFFFFFF, the rest truncate.
__device__ __forceinline__ uint32_t sat_u32(float x)
{
    if (!(x >= 1.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

// genMortonKeysKernel (RayBufferKernels.cu:139-171): a = (o - lo) / (hi -
// lo) and b = (normalize(d) + 1) / 2, quantised to 2^24 and 2^21 steps; bit i
// of coordinate k goes to key bit k + 6 i (collectBits :57-62).  Here with
// IEEE division and square root (the reference compiles with fast math).
// Key word w holds key bits 64 w .. 64 w + 63.
__global__ __launch_bounds__(256) void k_ray_keys(const float4* __restrict__ rays, uint32_t r,
                                                  const float* __restrict__ box,
                                                  uint64_t* __restrict__ k0,
                                                  uint64_t* __restrict__ k1,
                                                  uint64_t* __restrict__ k2)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= r) return;
    const float4 o = rays[2ull * i];
    const float4 d = rays[2ull * i + 1];
    const float ax = (o.x - box[0]) / (box[3] - box[0]);
    const float ay = (o.y - box[1]) / (box[4] - box[1]);
    const float az = (o.z - box[2]) / (box[5] - box[2]);
    const float len = sqrtf((d.x * d.x + d.y * d.y) + d.z * d.z);
    const float rl = len != 0.0f ? 1.0f / len : 0.0f;
    const float bx = (d.x * rl + 1.0f) * 0.5f;
    const float by = (d.y * rl + 1.0f) * 0.5f;
    const float bz = (d.z * rl + 1.0f) * 0.5f;
    const uint32_t q[6] = {sat_u32(ax * 256.0f * 65536.0f), sat_u32(ay * 256.0f * 65536.0f),
                           sat_u32(az * 256.0f * 65536.0f), sat_u32(bx * 32.0f * 65536.0f),
                           sat_u32(by * 32.0f * 65536.0f), sat_u32(bz * 32.0f * 65536.0f)};
    uint64_t w[3] = {0ull, 0ull, 0ull};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
#pragma unroll
        for (int b = 0; b < 32; ++b) {
            const int pos = k + 6 * b;
            w[pos >> 6] |= (uint64_t)((q[k] >> b) & 1u) << (pos & 63);
        }
    }
    k0[i] = w[0];
    k1[i] = w[1];
    k2[i] = w[2];
}

// key[i] = src[perm[i]].  Permutation entries are in range unless a bounded
// wait of the sort gave up (reported at the next synchronisation); an entry
// out of range then reads slot i instead, so every access stays in bounds.
__global__ __launch_bounds__(256) void k_gather_u64(const uint64_t* __restrict__ src,
                                                    const uint32_t* __restrict__ perm, uint32_t r,
                                                    uint64_t* __restrict__ dst)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= r) return;
    const uint32_t p = perm[i];
    dst[i] = src[p < r ? p : i];
}

// reorderRaysKernel (RayBufferKernels.cu:175-196): slot i takes the ray of
// its old slot; slot_to_id composes with the input's (identity when NULL).
__global__ __launch_bounds__(256) void k_reorder_rays(const float4* __restrict__ rays,
                                                      const uint32_t* __restrict__ perm,
                                                      const int32_t* __restrict__ in_slot_to_id,
                                                      uint32_t r, float4* __restrict__ out,
                                                      int32_t* __restrict__ slot_to_id,
                                                      int32_t* __restrict__ id_to_slot)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= r) return;
    uint32_t old = perm[i];
    if (old >= r) old = i;  // only after a sort that gave up (see k_gather_u64)
    const int32_t id = in_slot_to_id ? in_slot_to_id[old] : (int32_t)old;
    out[2ull * i] = rays[2ull * old];
    out[2ull * i + 1] = rays[2ull * old + 1];
    slot_to_id[i] = id;
    if (id_to_slot && (uint32_t)id < r) id_to_slot[id] = (int32_t)i;
}

// results[slot_to_id[s]] = sorted[s] (int4 per ray); ids out of range are
// skipped
__global__ __launch_bounds__(256) void k_unsort_results(const int4* __restrict__ sorted,
                                                        const int32_t* __restrict__ slot_to_id,
                                                        uint32_t r, int4* __restrict__ out)
{
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= r) return;
    const int32_t id = slot_to_id[s];
    if ((uint32_t)id < r) out[id] = sorted[s];
}

}  // namespace

atr_status run_ray_keys(atr_context* ctx, const float4* rays, uint32_t r, uint64_t* keys,
                        float* d_box)
{
    float* partial = nullptr;
    const uint32_t blocks = std::max(std::min<uint32_t>(div_up(r, 256 * 8), kBoxBlocks), 1u);
    ATR_TRY(dalloc_t(ctx, 6ull * blocks, &partial));
    ATR_TRY(launch(ctx, "ray_box", (double)r * 32.0, [&] {
        k_ray_box<<<blocks, 256, 0, ctx->stream>>>(rays, r, partial);
    }));
    ATR_TRY(launch(ctx, "ray_box_final", blocks * 24.0, [&] {
        k_ray_box_final<<<1, 256, 0, ctx->stream>>>(partial, blocks, d_box);
    }));
    atr_status s = launch(ctx, "ray_keys", (double)r * (32.0 + 24.0), [&] {
        k_ray_keys<<<std::max(div_up(r, 256), 1u), 256, 0, ctx->stream>>>(
            rays, r, d_box, keys, keys + r, keys + 2ull * r);
    });
    dfree(ctx, partial);
    return s;
}

// Passes of 8-bit digits per key word: the origin coordinates reach bit 24
// and the direction coordinates bit 21 (both in [0, 1]), so key bits above
// 146 are zero and the top word needs its low 24 bits only.
static const int kRayWordPasses[3] = {8, 8, 3};

atr_status run_ray_sort(atr_context* ctx, const float4* rays, uint32_t r,
                        const int32_t* in_slot_to_id, float4* out, int32_t* slot_to_id,
                        int32_t* id_to_slot)
{
    uint64_t* keys = nullptr;
    uint64_t* kalt = nullptr;
    uint32_t* perm[2] = {nullptr, nullptr};
    float* box = nullptr;
    atr_status s = dalloc_t(ctx, 3ull * r, &keys);
    if (s == ATR_OK) s = dalloc_t(ctx, r, &kalt);
    if (s == ATR_OK) s = dalloc_t(ctx, r, &perm[0]);
    if (s == ATR_OK) s = dalloc_t(ctx, r, &perm[1]);
    if (s == ATR_OK) s = dalloc_t(ctx, 8, &box);
    if (s == ATR_OK) s = run_ray_keys(ctx, rays, r, keys, box);
    // the sorted permutation lives in perm[cur]; word w's keys are gathered
    // through it into the scratch word, then sorted with it
    int cur = 0;
    bool first = true;
    for (int w = 0; w < 3 && s == ATR_OK; ++w) {
        uint64_t* kw = keys + (size_t)w * r;
        uint64_t* ksrc = kw;
        uint64_t* kdst = kalt;
        if (!first) {
            s = launch(ctx, "ray_key_gather", (double)r * 20.0, [&] {
                k_gather_u64<<<std::max(div_up(r, 256), 1u), 256, 0, ctx->stream>>>(
                    kw, perm[cur], r, kalt);
            });
            ksrc = kalt;
            kdst = kw;
        }
        bool in_alt = false;
        if (s == ATR_OK)
            s = run_sort(ctx, 64, kRayWordPasses[w], ksrc, perm[cur], r, kdst, perm[1 - cur],
                         first, nullptr, 0, nullptr, &in_alt);
        if (in_alt) cur = 1 - cur;
        first = false;
    }
    if (s == ATR_OK)
        s = launch(ctx, "reorder_rays", (double)r * 80.0, [&] {
            k_reorder_rays<<<std::max(div_up(r, 256), 1u), 256, 0, ctx->stream>>>(
                rays, perm[cur], in_slot_to_id, r, out, slot_to_id, id_to_slot);
        });
    dfree(ctx, keys);
    dfree(ctx, kalt);
    dfree(ctx, perm[0]);
    dfree(ctx, perm[1]);
    dfree(ctx, box);
    return s;
}

atr_status run_unsort_results(atr_context* ctx, const int4* sorted, const int32_t* slot_to_id,
                              uint32_t r, int4* out)
{
    return launch(ctx, "unsort_results", (double)r * 36.0, [&] {
        k_unsort_results<<<std::max(div_up(r, 256), 1u), 256, 0, ctx->stream>>>(sorted, slot_to_id,
                                                                                r, out);
    });
}

}  // namespace atr
