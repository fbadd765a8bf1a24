// Differentiable hits on gfx950: the adjoint of the closest hit (t, u, v)
// with respect to the hit triangle's vertices, scatter-added over a fixed
// topology, plus the depth loss and vertex step of the refit-driven
// optimisation workload.  No reference counterpart (the reference's tracer is
// not differentiable).
//
// A hit satisfies o + t d = (1-u-v) v0 + u v1 + v v2, i.e. M x = o - v0 with
// M = [a | e1 | e2], a = -d, e1 = v1 - v0, e2 = v2 - v0 and x = (t, u, v).
// Moving the vertices by dv_i changes x by dx = -M^{-1} (sum_i b_i dv_i) with
// b = (1-u-v, u, v).  For an upstream gradient g = dL/dx this gives
//   dL/dv_i = -b_i w,   w = M^{-T} g = (g_t (e1 x e2) + g_u (e2 x a) + g_v (a x e1)) / det M,
// det M = a . (e1 x e2).
#include "device.hpp"
#include "internal.hpp"

namespace atr {

namespace {

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 cross(V3 a, V3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

__global__ __launch_bounds__(256) void k_hit_adjoint(
    const float4* __restrict__ v, uint32_t n_tris, const float4* __restrict__ rays,
    const int4* __restrict__ hits, uint32_t R, const float* __restrict__ g_tuv,
    const uint32_t* __restrict__ index, uint32_t n_grad, float* grad)
{
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    if (r >= R) return;
    const int4 h = hits[r];
    if (h.x < 0 || (uint32_t)h.x >= n_tris) return;
    const float gt = g_tuv[3ull * r], gu = g_tuv[3ull * r + 1], gv = g_tuv[3ull * r + 2];
    if (gt == 0.0f && gu == 0.0f && gv == 0.0f) return;
    const float u = __int_as_float(h.z), w2 = __int_as_float(h.w);
    const uint32_t t = (uint32_t)h.x;
    const float4 p0 = v[3ull * t], p1 = v[3ull * t + 1], p2 = v[3ull * t + 2];
    const float4 d4 = rays[2ull * r + 1];
    const V3 a = {-d4.x, -d4.y, -d4.z};
    const V3 v0 = {p0.x, p0.y, p0.z};
    const V3 e1 = sub({p1.x, p1.y, p1.z}, v0);
    const V3 e2 = sub({p2.x, p2.y, p2.z}, v0);
    const V3 n = cross(e1, e2);
    const float det = dot(a, n);
    if (det == 0.0f) return;
    const V3 c1 = cross(e2, a);
    const V3 c2 = cross(a, e1);
    const float inv = 1.0f / det;
    const V3 w = {(gt * n.x + gu * c1.x + gv * c2.x) * inv, (gt * n.y + gu * c1.y + gv * c2.y) * inv,
                  (gt * n.z + gu * c1.z + gv * c2.z) * inv};
    const float b[3] = {1.0f - u - w2, u, w2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const uint32_t slot = index ? index[3ull * t + k] : (uint32_t)(3u * t + k);
        if (slot >= n_grad) continue;
        unsafeAtomicAdd(&grad[3ull * slot + 0], -b[k] * w.x);
        unsafeAtomicAdd(&grad[3ull * slot + 1], -b[k] * w.y);
        unsafeAtomicAdd(&grad[3ull * slot + 2], -b[k] * w.z);
    }
}

// Loss sum (t - t_target)^2 over rays that hit in both traces (and, when
// same_tri, hit the same triangle in both); per-ray upstream gradient
// (2 (t - t_target), 0, 0), zero elsewhere.
__global__ __launch_bounds__(256) void k_depth_loss(const int4* __restrict__ hits,
                                                    const int4* __restrict__ target, uint32_t R,
                                                    int same_tri, float* __restrict__ g_tuv,
                                                    double* loss)
{
    __shared__ double s_part[4];
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    double l = 0.0;
    if (r < R) {
        const int4 h = hits[r];
        const int4 tg = target[r];
        float g = 0.0f;
        if (h.x >= 0 && tg.x >= 0 && (!same_tri || h.x == tg.x)) {
            const float d = __int_as_float(h.y) - __int_as_float(tg.y);
            g = 2.0f * d;
            l = (double)d * (double)d;
        }
        g_tuv[3ull * r] = g;
        g_tuv[3ull * r + 1] = 0.0f;
        g_tuv[3ull * r + 2] = 0.0f;
    }
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) unsafeAtomicAdd(loss, (s_part[0] + s_part[1]) + (s_part[2] + s_part[3]));
}

// Gradient step on indexed mesh vertices: each vertex moves by -lr * grad,
// shortened to length max_step when it is longer (max_step <= 0: no limit);
// the gradient is cleared for the next accumulation.  Then the triangle soup
// is gathered from the mesh.
__global__ __launch_bounds__(256) void k_vertex_step(float4* mesh, uint32_t V, float* grad,
                                                     float lr, float max_step)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= V) return;
    float4 p = mesh[i];
    float sx = lr * grad[3ull * i];
    float sy = lr * grad[3ull * i + 1];
    float sz = lr * grad[3ull * i + 2];
    if (max_step > 0.0f) {
        const float len = sqrtf(sx * sx + sy * sy + sz * sz);
        if (len > max_step) {
            const float k = max_step / len;
            sx *= k;
            sy *= k;
            sz *= k;
        }
    }
    p.x -= sx;
    p.y -= sy;
    p.z -= sz;
    mesh[i] = p;
    grad[3ull * i] = 0.0f;
    grad[3ull * i + 1] = 0.0f;
    grad[3ull * i + 2] = 0.0f;
}

__global__ __launch_bounds__(256) void k_gather(const float4* __restrict__ mesh, uint32_t V,
                                                const uint32_t* __restrict__ index, uint64_t m,
                                                float4* __restrict__ soup)
{
    const uint64_t j = blockIdx.x * 256ull + threadIdx.x;
    if (j >= m) return;
    const uint32_t k = index[j];
    const float4 p = k < V ? mesh[k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    soup[j] = make_float4(p.x, p.y, p.z, 0.0f);
}

}  // namespace

atr_status run_hit_adjoint(atr_context* ctx, const float4* v, uint32_t n_tris,
                           const float4* rays, const int4* hits, uint32_t R, const float* g_tuv,
                           const uint32_t* index, uint32_t n_grad, float* grad)
{
    return launch(ctx, "hit_adjoint", (double)R * 76.0, [&] {
        k_hit_adjoint<<<std::max(div_up(R, 256), 1u), 256, 0, ctx->stream>>>(
            v, n_tris, rays, hits, R, g_tuv, index, n_grad, grad);
    });
}

atr_status run_depth_loss(atr_context* ctx, const int4* hits, const int4* target, uint32_t R,
                          int same_tri, float* g_tuv, double* loss)
{
    return launch(ctx, "depth_loss", (double)R * 44.0, [&] {
        k_depth_loss<<<std::max(div_up(R, 256), 1u), 256, 0, ctx->stream>>>(hits, target, R,
                                                                            same_tri, g_tuv, loss);
    });
}

atr_status run_mesh_step(atr_context* ctx, float4* mesh, uint32_t V, float* grad, float lr,
                         float max_step, const uint32_t* index, uint32_t n_tris, float4* soup)
{
    if (grad) {
        ATR_TRY(launch(ctx, "vertex_step", (double)V * 40.0, [&] {
            k_vertex_step<<<std::max(div_up(V, 256), 1u), 256, 0, ctx->stream>>>(mesh, V, grad,
                                                                                 lr, max_step);
        }));
    }
    if (!index) return ATR_OK;
    const uint64_t m = 3ull * n_tris;
    return launch(ctx, "gather", (double)m * 36.0, [&] {
        k_gather<<<std::max(div_up(m, 256), 1u), 256, 0, ctx->stream>>>(mesh, V, index, m, soup);
    });
}

}  // namespace atr
