// Closest-hit / any-hit traversal of the collapsed BVH on gfx950, with
// Moller-Trumbore intersection (Intersect::RayTriangle,
// gpu-ray-traversal/src/rt/Util.cpp:50-94) and the children ordering of
// BVH::traceRecursive (gpu-ray-traversal/src/rt/bvh/BVH.cpp:133-159): when both
// children are hit, the one with the smaller entry distance goes first.
// Ray generation for the primary / diffuse workloads.
#include "device.hpp"
#include "internal.hpp"

namespace atr {

namespace {

constexpr float kOoeps = 8.271806125530277e-25f;  // 2^-80
constexpr float kBoxEps = 1.52587890625e-05f;     // 2^-16
constexpr int kDone = (int)0x80000000;

struct Ray {
    float ox, oy, oz, tmin, dx, dy, dz, tmax;
};

__device__ __forceinline__ float safe_inv(float d)
{
    return 1.0f / (fabsf(d) > kOoeps ? d : copysignf(kOoeps, d));
}

// Slab test with conservative widening (entry shrunk and exit grown by 2^-16
// relative) so rounding never culls a box holding a triangle hit.
__device__ __forceinline__ bool slab(float lox, float hix, float loy, float hiy, float loz,
                                     float hiz, float ox, float oy, float oz, float ix, float iy,
                                     float iz, float tmin, float tmax, float& tnear)
{
    const float x0 = (lox - ox) * ix, x1 = (hix - ox) * ix;
    const float y0 = (loy - oy) * iy, y1 = (hiy - oy) * iy;
    const float z0 = (loz - oz) * iz, z1 = (hiz - oz) * iz;
    float lo = fmaxf(tmin, fminf(x0, x1));
    lo = fmaxf(lo, fminf(y0, y1));
    lo = fmaxf(lo, fminf(z0, z1));
    float hi = fminf(tmax, fmaxf(x0, x1));
    hi = fminf(hi, fmaxf(y0, y1));
    hi = fminf(hi, fmaxf(z0, z1));
    tnear = lo;
    return lo - fabsf(lo) * kBoxEps <= hi + fabsf(hi) * kBoxEps;
}

// Moller-Trumbore with precomputed edges e1 = v1 - v0, e2 = v2 - v0 (the same
// fp32 differences the reference forms per test).  dot(a,b) = (ax*bx + ay*by) + az*bz.
__device__ __forceinline__ bool intersect(const float4& v0, const float4& e1, const float4& e2,
                                          float ox, float oy, float oz, float dx, float dy,
                                          float dz, float tmin, float tmax, float& t, float& u,
                                          float& v)
{
    const float px = dy * e2.z - dz * e2.y;
    const float py = dz * e2.x - dx * e2.z;
    const float pz = dx * e2.y - dy * e2.x;
    const float det = e1.x * px + e1.y * py + e1.z * pz;
    const float tx = ox - v0.x, ty = oy - v0.y, tz = oz - v0.z;
    float uu = tx * px + ty * py + tz * pz;
    const float qx = ty * e1.z - tz * e1.y;
    const float qy = tz * e1.x - tx * e1.z;
    const float qz = tx * e1.y - ty * e1.x;
    float vv = dx * qx + dy * qy + dz * qz;
    if (det > 0.0f) {
        if (uu < 0.0f || uu > det) return false;
        if (vv < 0.0f || uu + vv > det) return false;
    } else if (det < 0.0f) {
        if (uu > 0.0f || uu < det) return false;
        if (vv > 0.0f || uu + vv < det) return false;
    } else {
        return false;
    }
    const float inv = 1.0f / det;
    const float tt = (e2.x * qx + e2.y * qy + e2.z * qz) * inv;
    if (!(tt > tmin && tt < tmax)) return false;
    t = tt;
    u = uu * inv;
    v = vv * inv;
    return true;
}

// One thread per ray.  `STACK` entries of private stack; the host only uses
// this kernel when the tree height fits, otherwise k_trace_deep.
template <bool ANY, int STACK>
__global__ __launch_bounds__(256) void k_trace(const float4* __restrict__ rays, uint32_t R,
                                               const float4* __restrict__ nodes,
                                               const float4* __restrict__ tris,
                                               const int32_t* __restrict__ tri_id,
                                               int4* __restrict__ results, int* __restrict__ gstack,
                                               int gdepth)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= R) return;
    const float4 r0 = rays[2ull * i];
    const float4 r1 = rays[2ull * i + 1];
    const float ox = r0.x, oy = r0.y, oz = r0.z, tmin = r0.w;
    const float dx = r1.x, dy = r1.y, dz = r1.z;
    float t = r1.w;
    const float ix = safe_inv(dx), iy = safe_inv(dy), iz = safe_inv(dz);
    int id = -1;
    float bu = 0.0f, bv = 0.0f;
    int stack[STACK > 0 ? STACK : 1];
    int* gs = gstack ? gstack + (size_t)i * (size_t)gdepth : nullptr;
    int sp = 0;
    int node = t > tmin ? 0 : kDone;
    while (node != kDone) {
        while (node >= 0) {
            const float4 n0 = nodes[4ull * node + 0];
            const float4 n1 = nodes[4ull * node + 1];
            const float4 n2 = nodes[4ull * node + 2];
            const float4 n3 = nodes[4ull * node + 3];
            float t0, t1;
            const bool h0 = slab(n0.x, n0.y, n0.z, n0.w, n2.x, n2.y, ox, oy, oz, ix, iy, iz, tmin,
                                 t, t0);
            const bool h1 = slab(n1.x, n1.y, n1.z, n1.w, n2.z, n2.w, ox, oy, oz, ix, iy, iz, tmin,
                                 t, t1);
            const int c0 = __float_as_int(n3.x);
            const int c1 = __float_as_int(n3.y);
            if (h0 && h1) {
                const bool swap = t0 > t1;
                const int nearc = swap ? c1 : c0;
                const int farc = swap ? c0 : c1;
                if (STACK > 0)
                    stack[sp++] = farc;
                else
                    gs[sp++] = farc;
                node = nearc;
            } else if (h0) {
                node = c0;
            } else if (h1) {
                node = c1;
            } else {
                node = sp > 0 ? (STACK > 0 ? stack[--sp] : gs[--sp]) : kDone;
            }
        }
        if (node == kDone) break;
        // leaf: slots ~node, ~node+1, ... up to the one tagged last
        bool stop = false;
        for (int k = ~node;; ++k) {
            const int tid = tri_id[k];
            const float4 v0 = tris[3ull * k];
            const float4 e1 = tris[3ull * k + 1];
            const float4 e2 = tris[3ull * k + 2];
            float tt, uu, vv;
            if (intersect(v0, e1, e2, ox, oy, oz, dx, dy, dz, tmin, t, tt, uu, vv)) {
                t = tt;
                bu = uu;
                bv = vv;
                id = tid & 0x7fffffff;
                if (ANY) {
                    stop = true;
                    break;
                }
            }
            if (tid < 0) break;
        }
        if (stop) break;
        node = sp > 0 ? (STACK > 0 ? stack[--sp] : gs[--sp]) : kDone;
    }
    results[i] = make_int4(id, __float_as_int(t), __float_as_int(bu), __float_as_int(bv));
}

// Persistent traversal with dynamic ray fetch (the scheme of the reference's
// kepler_dynamic_fetch kernel): each wave takes rays (from a RayQueue, see
// device.hpp) as soon as a quarter of its lanes are idle, so lanes whose ray
// missed early or finished early do not idle while the rest of the wave
// traverses.  Per ray the traversal is exactly k_trace's (same slab test,
// near child first, far child pushed, LIFO pops), so every ray gets the same
// result.  The most recent stack entry lives in a register (a pop returns it
// at once and the refill overlaps the next node fetch); the entries below it
// are an LdsStack: the newest 16 per lane in LDS, up to 48 more in private
// memory.  The host uses this kernel only when the tree height is below 64,
// which bounds the stack.
constexpr int kFetchRing = 16;
template <bool ANY>
__global__ __launch_bounds__(256) void k_trace_fetch(const float4* __restrict__ rays, uint32_t R,
                                                     const float4* __restrict__ nodes,
                                                     const float4* __restrict__ tris,
                                                     const int32_t* __restrict__ tri_id,
                                                     int4* __restrict__ results,
                                                     uint32_t* __restrict__ counter,
                                                     uint32_t* __restrict__ err)
{
    __shared__ uint32_t s_ring[kFetchRing * 256];
    LdsStack<kFetchRing, 48, 256> stack(s_ring + threadIdx.x);
    const uint64_t lt = lanemask_lt();
    bool active = false;
    RayQueue q(counter, R);
    uint32_t i = 0;
    float ox = 0.0f, oy = 0.0f, oz = 0.0f, tmin = 0.0f, dx = 0.0f, dy = 0.0f, dz = 0.0f;
    float t = 0.0f, ix = 0.0f, iy = 0.0f, iz = 0.0f, bu = 0.0f, bv = 0.0f;
    int id = -1, node = kDone, sp = 0, top = 0;
    while (true) {
        const uint64_t idle = __ballot(!active);
        if (!q.exhausted && __popcll(idle) >= 16) {
            const int leader = __ffsll((unsigned long long)idle) - 1;
            uint32_t first = 0;
            const uint32_t take = q.take((uint32_t)__popcll(idle), leader, first);
            const uint32_t rank = (uint32_t)__popcll(idle & lt);
            if (!active && rank < take) {
                const uint32_t k = q.ray_of(first + rank);
                if (k < R) {
                    i = k;
                    const float4 r0 = rays[2ull * i];
                    const float4 r1 = rays[2ull * i + 1];
                    ox = r0.x;
                    oy = r0.y;
                    oz = r0.z;
                    tmin = r0.w;
                    dx = r1.x;
                    dy = r1.y;
                    dz = r1.z;
                    t = r1.w;
                    ix = safe_inv(dx);
                    iy = safe_inv(dy);
                    iz = safe_inv(dz);
                    id = -1;
                    bu = 0.0f;
                    bv = 0.0f;
                    sp = 0;
                    stack.n = 0;
                    node = t > tmin ? 0 : kDone;
                    active = true;
                }
            }
        }
        if (__ballot(active) == 0ull) {
            if (q.exhausted) break;
            continue;
        }
        // internal nodes until every active lane is at a leaf or done
        while (__ballot(active && node >= 0) != 0ull) {
            if (active && node >= 0) {
                const float4 n0 = nodes[4ull * node + 0];
                const float4 n1 = nodes[4ull * node + 1];
                const float4 n2 = nodes[4ull * node + 2];
                const float4 n3 = nodes[4ull * node + 3];
                float t0, t1;
                const bool h0 = slab(n0.x, n0.y, n0.z, n0.w, n2.x, n2.y, ox, oy, oz, ix, iy, iz,
                                     tmin, t, t0);
                const bool h1 = slab(n1.x, n1.y, n1.z, n1.w, n2.z, n2.w, ox, oy, oz, ix, iy, iz,
                                     tmin, t, t1);
                const int c0 = __float_as_int(n3.x);
                const int c1 = __float_as_int(n3.y);
                if (h0 && h1) {
                    const bool swap = t0 > t1;
                    if (sp > 0 && !stack.push((uint32_t)top)) {
                        atomicOr(err, 2u);
                        node = kDone;
                        id = -1;
                    } else {
                        top = swap ? c0 : c1;
                        ++sp;
                        node = swap ? c1 : c0;
                    }
                } else if (h0) {
                    node = c0;
                } else if (h1) {
                    node = c1;
                } else if (sp > 0) {
                    node = top;
                    --sp;
                    if (sp > 0) top = (int)stack.pop();
                } else {
                    node = kDone;
                }
            }
        }
        if (active && node != kDone) {
            // leaf: slots ~node, ~node+1, ... up to the one tagged last
            bool stop = false;
            for (int k = ~node;; ++k) {
                const int tid = tri_id[k];
                const float4 v0 = tris[3ull * k];
                const float4 e1 = tris[3ull * k + 1];
                const float4 e2 = tris[3ull * k + 2];
                float tt, uu, vv;
                if (intersect(v0, e1, e2, ox, oy, oz, dx, dy, dz, tmin, t, tt, uu, vv)) {
                    t = tt;
                    bu = uu;
                    bv = vv;
                    id = tid & 0x7fffffff;
                    if (ANY) {
                        stop = true;
                        break;
                    }
                }
                if (tid < 0) break;
            }
            if (stop || sp == 0) {
                node = kDone;
            } else {
                node = top;
                --sp;
                if (sp > 0) top = (int)stack.pop();
            }
        }
        if (active && node == kDone) {
            results[i] = make_int4(id, __float_as_int(t), __float_as_int(bu), __float_as_int(bv));
            active = false;
        }
    }
}

__global__ void k_zero_words(uint32_t* p) { p[threadIdx.x] = 0u; }

// Closest-hit traversal of a batch's binary trees (the forest): ray r belongs
// to mesh r / rays_per_mesh and starts at that mesh's root (internal slot
// seg[m]); leaves test their triangle (index into the batch's vertices).
// Same slab test, child order and Moller-Trumbore as k_trace, so hits equal
// the per-mesh collapsed traversal's except for exact ties.  A tree deeper
// than the stack stops the ray and raises bit 1 of `err`.
constexpr int kForestStack = 128;
__global__ __launch_bounds__(256) void k_trace_forest(
    const float4* __restrict__ rays, uint32_t R, uint32_t rays_per_mesh,
    const uint32_t* __restrict__ seg, int first_leaf, const int32_t* __restrict__ left,
    const int32_t* __restrict__ right, const int32_t* __restrict__ data,
    const float4* __restrict__ bbmin, const float4* __restrict__ bbmax,
    const float4* __restrict__ v, int4* __restrict__ results, uint32_t* err)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= R) return;
    const float4 r0 = rays[2ull * i];
    const float4 r1 = rays[2ull * i + 1];
    const float ox = r0.x, oy = r0.y, oz = r0.z, tmin = r0.w;
    const float dx = r1.x, dy = r1.y, dz = r1.z;
    float t = r1.w;
    const float ix = safe_inv(dx), iy = safe_inv(dy), iz = safe_inv(dz);
    int id = -1;
    float bu = 0.0f, bv = 0.0f;
    int stack[kForestStack];
    int sp = 0;
    int node = t > tmin ? (int)seg[i / rays_per_mesh] : -1;
    while (node >= 0) {
        if (node >= first_leaf) {
            const int tri = data[node];
            const float4 a = v[3ull * tri], b = v[3ull * tri + 1], c = v[3ull * tri + 2];
            const float4 e1 = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, 0.0f);
            const float4 e2 = make_float4(c.x - a.x, c.y - a.y, c.z - a.z, 0.0f);
            float tt, uu, vv;
            if (intersect(a, e1, e2, ox, oy, oz, dx, dy, dz, tmin, t, tt, uu, vv)) {
                t = tt;
                bu = uu;
                bv = vv;
                id = tri;
            }
            node = sp > 0 ? stack[--sp] : -1;
            continue;
        }
        const int c0 = left[node], c1 = right[node];
        const float4 a0 = bbmin[c0], a1 = bbmax[c0], b0 = bbmin[c1], b1 = bbmax[c1];
        float t0, t1;
        const bool h0 = slab(a0.x, a1.x, a0.y, a1.y, a0.z, a1.z, ox, oy, oz, ix, iy, iz, tmin, t,
                             t0);
        const bool h1 = slab(b0.x, b1.x, b0.y, b1.y, b0.z, b1.z, ox, oy, oz, ix, iy, iz, tmin, t,
                             t1);
        if (h0 && h1) {
            const bool swap = t0 > t1;
            if (sp == kForestStack) {
                atomicOr(err, 2u);
                break;
            }
            stack[sp++] = swap ? c0 : c1;
            node = swap ? c1 : c0;
        } else if (h0) {
            node = c0;
        } else if (h1) {
            node = c1;
        } else {
            node = sp > 0 ? stack[--sp] : -1;
        }
    }
    results[i] = make_int4(id, __float_as_int(t), __float_as_int(bu), __float_as_int(bv));
}

// ------------------------------------------------------------------ ray gen
__global__ __launch_bounds__(256) void k_rays_pinhole(float ex, float ey, float ez, float rx,
                                                      float ry, float rz, float ux, float uy,
                                                      float uz, float fx, float fy, float fz,
                                                      float tan_half, uint32_t w, uint32_t h,
                                                      float tmax, float4* rays)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= w * h) return;
    const uint32_t x = i % w, y = i / w;
    const float sx = (2.0f * ((float)x + 0.5f) / (float)w - 1.0f) * tan_half;
    const float sy = (2.0f * ((float)y + 0.5f) / (float)h - 1.0f) * tan_half;
    float dx = sx * rx + sy * ux + fx;
    float dy = sx * ry + sy * uy + fy;
    float dz = sx * rz + sy * uz + fz;
    const float len = sqrtf(dx * dx + dy * dy + dz * dz);
    dx /= len;
    dy /= len;
    dz /= len;
    rays[2ull * i] = make_float4(ex, ey, ez, 0.0f);
    rays[2ull * i + 1] = make_float4(dx, dy, dz, tmax);
}

// Bob Jenkins' 96-bit mix (public domain lookup2 mix), as used by the
// reference ray generator.
__device__ __forceinline__ void jenkins_mix(uint32_t& a, uint32_t& b, uint32_t& c)
{
    a -= b; a -= c; a ^= (c >> 13);
    b -= c; b -= a; b ^= (a << 8);
    c -= a; c -= b; c ^= (b >> 13);
    a -= b; a -= c; a ^= (c >> 12);
    b -= c; b -= a; b ^= (a << 16);
    c -= a; c -= b; c ^= (b >> 5);
    a -= b; a -= c; a ^= (c >> 3);
    b -= c; b -= a; b ^= (a << 10);
    c -= a; c -= b; c ^= (b >> 15);
}

// normalize() of the reference's vector library (framework/base/Math.hpp:
// 139-141): v * rcp(sqrt((x*x + y*y) + z*z)) with rcp(0) = 0, so a zero
// vector stays zero.
__device__ __forceinline__ void normalize3(float& x, float& y, float& z)
{
    const float l = sqrtf((x * x + y * y) + z * z);
    const float r = l != 0.0f ? 1.0f / l : 0.0f;
    x *= r;
    y *= r;
    z *= r;
}

// One cosine-weighted ray per primary ray (rayGenAOKernel with numSamples = 1,
// gpu-ray-traversal/src/rt/ray/RayGenKernels.cu:114-221): origin backed off by
// 1e-4 along the primary ray, the hit triangle's normal (Scene.cpp:74:
// normalize(cross(v1 - v0, v2 - v0))) flipped towards the ray, frame rotated
// by a hashed angle, Halton (2, 3) sample 1 warped to the hemisphere.  A
// primary miss gives a degenerate ray (tmax = -1).
__global__ __launch_bounds__(256) void k_rays_diffuse(const float4* __restrict__ v,
                                                      const float4* __restrict__ prim,
                                                      const int4* __restrict__ res, uint32_t R,
                                                      uint32_t seed, float tmax, float4* rays)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= R) return;
    const float4 r0 = prim[2ull * i];
    const float4 r1 = prim[2ull * i + 1];
    const int4 hit = res[i];
    const float ht = __int_as_float(hit.y);
    const float back = fmaxf(ht - 1.0e-4f, 0.0f);
    const float ox = r0.x + r1.x * back, oy = r0.y + r1.y * back, oz = r0.z + r1.z * back;
    float nx = 1.0f, ny = 0.0f, nz = 0.0f;
    const int tri = hit.x;
    if (tri >= 0) {
        const float4 a = v[3ull * tri], b = v[3ull * tri + 1], c = v[3ull * tri + 2];
        const float e1x = b.x - a.x, e1y = b.y - a.y, e1z = b.z - a.z;
        const float e2x = c.x - a.x, e2y = c.y - a.y, e2z = c.z - a.z;
        nx = e1y * e2z - e1z * e2y;
        ny = e1z * e2x - e1x * e2z;
        nz = e1x * e2y - e1y * e2x;
        normalize3(nx, ny, nz);
    }
    if ((nx * r1.x + ny * r1.y) + nz * r1.z > 0.0f) {
        nx = -nx;
        ny = -ny;
        nz = -nz;
    }
    const float ax = fabsf(nx), ay = fabsf(ny), az = fabsf(nz);
    const float nm = fmaxf(fmaxf(ax, ay), az);
    float px = ny, py = -nx, pz = 0.0f;
    if (nm == az) {
        px = 0.0f;
        py = nz;
        pz = -ny;
    } else if (nm == ax) {
        px = -nz;
        py = 0.0f;
        pz = nx;
    }
    normalize3(px, py, pz);
    const float bx = ny * pz - nz * py, by = nz * px - nx * pz, bz = nx * py - ny * px;
    uint32_t ha = seed + i, hb = 0x9e3779b9u, hc = 0x9e3779b9u;
    jenkins_mix(ha, hb, hc);
    jenkins_mix(ha, hb, hc);
    const float kPi = 3.14159265358979323846f;
    const float angle = 2.0f * kPi * (float)hc * 2.3283064365386963e-10f;
    const float ca = cosf(angle), sa = sinf(angle);
    const float t0x = px * ca + bx * sa, t0y = py * ca + by * sa, t0z = pz * ca + bz * sa;
    const float t1x = px * -sa + bx * ca, t1y = py * -sa + by * ca, t1z = pz * -sa + bz * ca;
    // Halton sample index 1: base 2 -> 0.5, base 3 -> 1/3
    const float hx = 0.5f, hy = 1.0f / 3.0f;
    const float phi = 2.0f * kPi * hy;
    const float rr = sqrtf(hx);
    const float sx = rr * cosf(phi), sy = rr * sinf(phi);
    const float sz = sqrtf(1.0f - sx * sx - sy * sy);
    float dx = (sx * t0x + sy * t1x) + sz * nx;
    float dy = (sx * t0y + sy * t1y) + sz * ny;
    float dz = (sx * t0z + sy * t1z) + sz * nz;
    normalize3(dx, dy, dz);
    rays[2ull * i] = make_float4(ox, oy, oz, 0.0f);
    rays[2ull * i + 1] = make_float4(dx, dy, dz, tri >= 0 ? tmax : -1.0f);
}

}  // namespace

atr_status run_trace_forest(atr_context* ctx, const atr_tree* f, const uint32_t* d_seg,
                            const float4* verts, const float4* rays, uint32_t r,
                            uint32_t rays_per_mesh, int32_t* results)
{
    if (r == 0) return ATR_OK;
    return launch(ctx, "trace_forest", (double)r * 48.0, [&] {
        k_trace_forest<<<div_up(r, 256), 256, 0, ctx->stream>>>(
            rays, r, rays_per_mesh, d_seg, (int)f->first_leaf, f->left, f->right, f->data,
            f->bbmin, f->bbmax, verts, (int4*)results, ctx->d_status);
    });
}

atr_status run_trace(atr_context* ctx, atr_collapsed* c, const float4* rays, uint32_t r,
                     int any_hit, int32_t* results, int depth)
{
    if (r == 0) return ATR_OK;
    const uint32_t grid = div_up(r, 256);
    int4* out = (int4*)results;
    const double bytes = (double)r * 48.0;
    if (depth < 64) {
        // persistent waves: enough workgroups to fill every CU, at most one
        // per 256 rays
        if (ctx->cu_count == 0) {
            int cus = 0;
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) !=
                    hipSuccess ||
                cus <= 0)
                cus = 256;
            ctx->cu_count = cus;
        }
        const uint32_t pgrid = std::min<uint32_t>(grid, (uint32_t)ctx->cu_count * 8u);
        uint32_t* counter = nullptr;
        ATR_TRY(dalloc_t(ctx, kRayCounterWords, &counter));
        atr_status s = launch(ctx, "trace_zero", kRayCounterWords * 4.0, [&] {
            k_zero_words<<<1, kRayCounterWords, 0, ctx->stream>>>(counter);
        });
        if (s == ATR_OK)
            s = launch(ctx, any_hit ? "trace_any" : "trace", bytes, [&] {
                if (any_hit)
                    k_trace_fetch<true><<<pgrid, 256, 0, ctx->stream>>>(
                        rays, r, c->tnodes, c->ttris, c->ttri_id, out, counter, ctx->d_status);
                else
                    k_trace_fetch<false><<<pgrid, 256, 0, ctx->stream>>>(
                        rays, r, c->tnodes, c->ttris, c->ttri_id, out, counter, ctx->d_status);
            });
        dfree(ctx, counter);
        return s;
    }
    int* gstack = nullptr;
    const int gdepth = depth + 1;
    ATR_TRY(dalloc_t(ctx, (size_t)r * gdepth, &gstack));
    atr_status s = launch(ctx, any_hit ? "trace_any_deep" : "trace_deep", bytes, [&] {
        if (any_hit)
            k_trace<true, 0><<<grid, 256, 0, ctx->stream>>>(rays, r, c->tnodes, c->ttris,
                                                            c->ttri_id, out, gstack, gdepth);
        else
            k_trace<false, 0><<<grid, 256, 0, ctx->stream>>>(rays, r, c->tnodes, c->ttris,
                                                             c->ttri_id, out, gstack, gdepth);
    });
    dfree(ctx, gstack);
    return s;
}

atr_status run_rays_pinhole(atr_context* ctx, const float* eye, const float* basis,
                            float tan_half, uint32_t w, uint32_t h, float tmax, float4* rays)
{
    const uint64_t count = (uint64_t)w * h;
    return launch(ctx, "rays_pinhole", count * 32.0, [&] {
        k_rays_pinhole<<<div_up(count, 256), 256, 0, ctx->stream>>>(
            eye[0], eye[1], eye[2], basis[0], basis[1], basis[2], basis[3], basis[4], basis[5],
            basis[6], basis[7], basis[8], tan_half, w, h, tmax, rays);
    });
}

atr_status run_rays_diffuse(atr_context* ctx, const float4* verts, const float4* primary,
                            const int32_t* results, uint32_t r, uint32_t seed, float tmax,
                            float4* rays)
{
    return launch(ctx, "rays_diffuse", r * 96.0, [&] {
        k_rays_diffuse<<<div_up(r, 256), 256, 0, ctx->stream>>>(
            verts, primary, (const int4*)results, r, seed, tmax, rays);
    });
}

}  // namespace atr
