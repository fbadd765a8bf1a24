// CPU baseline of the reference's CPU path (test / benchmark infrastructure
// only): the split-BVH (SBVH) builder of
// gpu-ray-traversal/src/rt/bvh/SplitBVHBuilder.cpp:52-476 with the "GPU"
// platform of gpu-ray-traversal/src/rt/cuda/Renderer.cpp:71-72 (node and
// triangle SAH cost 1, leaves of 1..8 triangles), alpha = 1e-5
// (gpu-ray-traversal/src/rt/bvh/BVH.hpp:76), 128 spatial bins, maximum depth
// 64 and spatial-split depth 48 (SplitBVHBuilder.hpp:41-43); and a
// multi-threaded closest-hit tracer over it with the Moller-Trumbore and box
// semantics of BVH::traceRecursive / Intersect::RayTriangle / RayBox
// (gpu-ray-traversal/src/rt/bvh/BVH.cpp:67-163, src/rt/Util.cpp:34-94).
//
// Every node works on its own reference list, in the order the reference's
// shared reference stack holds it, so subtrees build independently: one
// thread gives the reference's tree, and several threads (a task per subtree
// near the root) give the same tree.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <future>
#include <thread>
#include <vector>

#include "oracle.h"

namespace {

struct Box {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const float* p)
    {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    void grow(const Box& b)
    {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    void intersect(const Box& b)
    {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::max(lo[a], b.lo[a]);
            hi[a] = std::min(hi[a], b.hi[a]);
        }
    }
    bool valid() const { return lo[0] <= hi[0] && lo[1] <= hi[1] && lo[2] <= hi[2]; }
    float area() const
    {
        if (!valid()) return 0.0f;
        const float x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
        return (x * y + y * z + z * x) * 2.0f;
    }
};

struct Ref {
    int tri;
    Box box;
};

constexpr int kBins = 128;
constexpr int kMaxDepth = 64;
constexpr int kMaxSpatialDepth = 48;
constexpr int kMinLeaf = 1;
constexpr int kMaxLeaf = 8;

struct Node {
    Box box;
    Node* kid[2] = {nullptr, nullptr};
    std::vector<int> tris;  // leaves only, in the order the leaf takes them
};

struct Builder {
    const float* v;  // float4[3n]
    float min_overlap = 0.0f;
    std::atomic<long long> duplicates{0};
    int par_depth = 0;

    const float* vert(int tri, int k) const { return v + ((size_t)tri * 3 + k) * 4; }

    static float tri_cost(int n) { return (float)n; }
    static float node_cost(int n) { return (float)n; }

    struct ObjSplit {
        float sah = FLT_MAX;
        int dim = 0;
        int num_left = 0;
        Box lb, rb;
    };
    struct SpatSplit {
        float sah = FLT_MAX;
        int dim = 0;
        float pos = 0.0f;
    };

    static void sort_dim(std::vector<Ref>& r, int dim)
    {
        std::sort(r.begin(), r.end(), [dim](const Ref& a, const Ref& b) {
            const float ca = a.box.lo[dim] + a.box.hi[dim];
            const float cb = b.box.lo[dim] + b.box.hi[dim];
            return ca < cb || (ca == cb && a.tri < b.tri);
        });
    }

    ObjSplit find_object(std::vector<Ref>& r, float node_sah)
    {
        ObjSplit s;
        float best_tie = FLT_MAX;
        const int n = (int)r.size();
        std::vector<Box> right(n > 1 ? n - 1 : 1);
        for (int dim = 0; dim < 3; ++dim) {
            sort_dim(r, dim);
            Box rb;
            for (int i = n - 1; i > 0; --i) {
                rb.grow(r[i].box);
                right[i - 1] = rb;
            }
            Box lb;
            for (int i = 1; i < n; ++i) {
                lb.grow(r[i - 1].box);
                const float sah = node_sah + lb.area() * tri_cost(i) +
                                  right[i - 1].area() * tri_cost(n - i);
                const float tie = (float)i * (float)i + (float)(n - i) * (float)(n - i);
                if (sah < s.sah || (sah == s.sah && tie < best_tie)) {
                    s.sah = sah;
                    s.dim = dim;
                    s.num_left = i;
                    s.lb = lb;
                    s.rb = right[i - 1];
                    best_tie = tie;
                }
            }
        }
        return s;
    }

    void split_ref(Ref& l, Ref& rr, const Ref& ref, int dim, float pos) const
    {
        l.tri = rr.tri = ref.tri;
        l.box = rr.box = Box();
        const float* p1 = vert(ref.tri, 2);
        for (int i = 0; i < 3; ++i) {
            const float* p0 = p1;
            p1 = vert(ref.tri, i);
            const float a = p0[dim], b = p1[dim];
            if (a <= pos) l.box.grow(p0);
            if (a >= pos) rr.box.grow(p0);
            if ((a < pos && b > pos) || (a > pos && b < pos)) {
                const float t = std::min(std::max((pos - a) / (b - a), 0.0f), 1.0f);
                float q[3];
                for (int k = 0; k < 3; ++k) q[k] = p0[k] + (p1[k] - p0[k]) * t;
                l.box.grow(q);
                rr.box.grow(q);
            }
        }
        l.box.hi[dim] = pos;
        rr.box.lo[dim] = pos;
        l.box.intersect(ref.box);
        rr.box.intersect(ref.box);
    }

    SpatSplit find_spatial(const std::vector<Ref>& r, const Box& nb, float node_sah)
    {
        struct Bin {
            Box box;
            int enter = 0, exit = 0;
        };
        std::vector<Bin> bins(3 * kBins);
        float origin[3], size[3], inv[3];
        for (int a = 0; a < 3; ++a) {
            origin[a] = nb.lo[a];
            size[a] = (nb.hi[a] - origin[a]) * (1.0f / (float)kBins);
            inv[a] = 1.0f / size[a];
        }
        for (const Ref& ref : r) {
            int first[3], last[3];
            for (int a = 0; a < 3; ++a) {
                first[a] = std::min(std::max((int)((ref.box.lo[a] - origin[a]) * inv[a]), 0),
                                    kBins - 1);
                last[a] = std::min(std::max((int)((ref.box.hi[a] - origin[a]) * inv[a]), first[a]),
                                   kBins - 1);
            }
            for (int a = 0; a < 3; ++a) {
                Ref cur = ref;
                for (int i = first[a]; i < last[a]; ++i) {
                    Ref lr, rr;
                    split_ref(lr, rr, cur, a, origin[a] + size[a] * (float)(i + 1));
                    bins[a * kBins + i].box.grow(lr.box);
                    cur = rr;
                }
                bins[a * kBins + last[a]].box.grow(cur.box);
                bins[a * kBins + first[a]].enter++;
                bins[a * kBins + last[a]].exit++;
            }
        }
        SpatSplit s;
        Box right[kBins - 1];
        for (int a = 0; a < 3; ++a) {
            Box rb;
            for (int i = kBins - 1; i > 0; --i) {
                rb.grow(bins[a * kBins + i].box);
                right[i - 1] = rb;
            }
            Box lb;
            int nl = 0, nr = (int)r.size();
            for (int i = 1; i < kBins; ++i) {
                lb.grow(bins[a * kBins + i - 1].box);
                nl += bins[a * kBins + i - 1].enter;
                nr -= bins[a * kBins + i - 1].exit;
                const float sah = node_sah + lb.area() * tri_cost(nl) +
                                  right[i - 1].area() * tri_cost(nr);
                if (sah < s.sah) {
                    s.sah = sah;
                    s.dim = a;
                    s.pos = origin[a] + size[a] * (float)i;
                }
            }
        }
        return s;
    }

    // performSpatialSplit: `r` becomes the left list, `right` the right list,
    // in the reference stack's order.
    void perform_spatial(std::vector<Ref>& r, std::vector<Ref>& right, Box& lbox, Box& rbox,
                         const SpatSplit& s)
    {
        std::vector<Ref>& refs = r;
        int left_end = 0;
        int right_start = (int)refs.size();
        lbox = rbox = Box();
        for (int i = left_end; i < right_start; ++i) {
            if (refs[i].box.hi[s.dim] <= s.pos) {
                lbox.grow(refs[i].box);
                std::swap(refs[i], refs[left_end++]);
            } else if (refs[i].box.lo[s.dim] >= s.pos) {
                rbox.grow(refs[i].box);
                std::swap(refs[i--], refs[--right_start]);
            }
        }
        while (left_end < right_start) {
            Ref lr, rr;
            split_ref(lr, rr, refs[left_end], s.dim, s.pos);
            Box lub = lbox, rub = rbox, ldb = lbox, rdb = rbox;
            lub.grow(refs[left_end].box);
            rub.grow(refs[left_end].box);
            ldb.grow(lr.box);
            rdb.grow(rr.box);
            const float lac = tri_cost(left_end);
            const float rac = tri_cost((int)refs.size() - right_start);
            const float lbc = tri_cost(left_end + 1);
            const float rbc = tri_cost((int)refs.size() - right_start + 1);
            const float ul = lub.area() * lbc + rbox.area() * rac;
            const float ur = lbox.area() * lac + rub.area() * rbc;
            const float du = ldb.area() * lbc + rdb.area() * rbc;
            const float m = std::min(ul, std::min(ur, du));
            if (m == ul) {
                lbox = lub;
                left_end++;
            } else if (m == ur) {
                rbox = rub;
                std::swap(refs[left_end], refs[--right_start]);
            } else {
                lbox = ldb;
                rbox = rdb;
                refs[left_end++] = lr;
                refs.push_back(rr);
            }
        }
        right.assign(refs.begin() + right_start, refs.end());
        refs.resize(left_end);
    }

    Node* build(std::vector<Ref> r, Box nb, int level)
    {
        // remove degenerates, top of the stack first with swap-removal
        for (int i = (int)r.size() - 1; i >= 0; --i) {
            const Box& b = r[i].box;
            const float sx = b.hi[0] - b.lo[0], sy = b.hi[1] - b.lo[1], sz = b.hi[2] - b.lo[2];
            const float mn = std::min(sx, std::min(sy, sz));
            const float mx = std::max(sx, std::max(sy, sz));
            if (mn < 0.0f || sx + sy + sz == mx) {
                r[i] = r.back();
                r.pop_back();
            }
        }
        const int n = (int)r.size();
        if (n <= kMinLeaf || level >= kMaxDepth) return leaf(r, nb);
        const float area = nb.area();
        const float leaf_sah = area * tri_cost(n);
        const float node_sah = area * node_cost(2);
        ObjSplit obj = find_object(r, node_sah);
        SpatSplit sp;
        if (level < kMaxSpatialDepth) {
            Box ov = obj.lb;
            ov.intersect(obj.rb);
            if (ov.area() >= min_overlap) sp = find_spatial(r, nb, node_sah);
        }
        const float m = std::min(leaf_sah, std::min(obj.sah, sp.sah));
        if (m == leaf_sah && n <= kMaxLeaf) return leaf(r, nb);
        std::vector<Ref> left, right;
        Box lb, rb;
        if (m == sp.sah) {
            left = r;
            perform_spatial(left, right, lb, rb, sp);
        }
        if (left.empty() || right.empty()) {
            sort_dim(r, obj.dim);
            left.assign(r.begin(), r.begin() + obj.num_left);
            right.assign(r.begin() + obj.num_left, r.end());
            lb = obj.lb;
            rb = obj.rb;
        }
        duplicates += (long long)left.size() + (long long)right.size() - n;
        std::vector<Ref>().swap(r);
        Node* node = new Node();
        node->box = nb;
        if (level < par_depth) {
            auto fr = std::async(std::launch::async,
                                 [&]() { return build(std::move(right), rb, level + 1); });
            node->kid[0] = build(std::move(left), lb, level + 1);
            node->kid[1] = fr.get();
        } else {
            // the reference builds the right child first; subtrees are
            // independent here, so the order does not change the tree
            node->kid[1] = build(std::move(right), rb, level + 1);
            node->kid[0] = build(std::move(left), lb, level + 1);
        }
        return node;
    }

    static Node* leaf(const std::vector<Ref>& r, const Box& nb)
    {
        Node* node = new Node();
        node->box = nb;
        for (int i = (int)r.size() - 1; i >= 0; --i) node->tris.push_back(r[i].tri);
        return node;
    }
};

struct Flat {
    std::vector<float> lo, hi;      // 3 per node
    std::vector<int32_t> kid;       // 2 per node; leaves: -1
    std::vector<int32_t> first;     // leaves: first triangle slot
    std::vector<int32_t> count;     // leaves: triangle count
    std::vector<int32_t> tri;       // triangle slots
};

// Pre-order flattening (left child first) plus the two SAH figures: the
// reference's top-down SAH (BVHNode.cpp:66-81: sum of area ratio x cost, inner
// nodes cost 2, leaves their triangle count) and the BVHRT convention
// (1.2 x inner areas + 1 x leaf areas x triangles, over the root area).
void flatten(Node* node, Flat& f, double root_area, double* sah_sbvh, double* sah_bvhrt,
             int* max_depth)
{
    std::vector<Node*> pre;
    std::vector<int> dep;
    std::vector<Node*> stack{node};
    std::vector<int> dstack{0};
    while (!stack.empty()) {
        Node* x = stack.back();
        stack.pop_back();
        const int d = dstack.back();
        dstack.pop_back();
        pre.push_back(x);
        dep.push_back(d);
        if (!x->tris.empty() || !x->kid[0]) continue;
        stack.push_back(x->kid[1]);
        dstack.push_back(d + 1);
        stack.push_back(x->kid[0]);
        dstack.push_back(d + 1);
    }
    const size_t m = pre.size();
    f.lo.resize(3 * m);
    f.hi.resize(3 * m);
    f.kid.assign(2 * m, -1);
    f.first.assign(m, 0);
    f.count.assign(m, 0);
    std::vector<std::pair<Node*, int>> idx;
    idx.reserve(m);
    for (size_t i = 0; i < m; ++i) idx.emplace_back(pre[i], (int)i);
    std::sort(idx.begin(), idx.end());
    auto find = [&](Node* x) {
        return std::lower_bound(idx.begin(), idx.end(), std::make_pair(x, -1))->second;
    };
    double si = 0.0, sl = 0.0, s_sbvh = 0.0;
    for (size_t i = 0; i < m; ++i) {
        Node* x = pre[i];
        for (int a = 0; a < 3; ++a) {
            f.lo[3 * i + a] = x->box.lo[a];
            f.hi[3 * i + a] = x->box.hi[a];
        }
        const double p = (double)x->box.area() / root_area;
        *max_depth = std::max(*max_depth, dep[i]);
        if (x->kid[0]) {
            f.kid[2 * i] = find(x->kid[0]);
            f.kid[2 * i + 1] = find(x->kid[1]);
            si += x->box.area();
            s_sbvh += p * 2.0;
        } else {
            f.first[i] = (int)f.tri.size();
            f.count[i] = (int)x->tris.size();
            for (int t : x->tris) f.tri.push_back(t);
            sl += (double)x->box.area() * (double)x->tris.size();
            s_sbvh += p * (double)x->tris.size();
        }
    }
    *sah_sbvh = s_sbvh;
    *sah_bvhrt = (1.2 * si + 1.0 * sl) / root_area;
}

void free_tree(Node* x)
{
    std::vector<Node*> st{x};
    while (!st.empty()) {
        Node* y = st.back();
        st.pop_back();
        if (y->kid[0]) st.push_back(y->kid[0]);
        if (y->kid[1]) st.push_back(y->kid[1]);
        delete y;
    }
}

struct SbvhHandle {
    Flat f;
    orc_sbvh_stats stats;
};

// Intersect::RayBox (Util.cpp:34-46): entry and exit of the slabs, by
// division as the reference computes them.
void ray_box(const float* lo, const float* hi, const float* o, const float* d, float* tn,
             float* tf)
{
    float a = -FLT_MAX, b = FLT_MAX;
    for (int k = 0; k < 3; ++k) {
        const float t0 = (lo[k] - o[k]) / d[k];
        const float t1 = (hi[k] - o[k]) / d[k];
        a = std::max(a, std::min(t0, t1));
        b = std::min(b, std::max(t0, t1));
    }
    *tn = a;
    *tf = b;
}

}  // namespace

extern "C" void* orc_sbvh_build(const float* v4, uint32_t n, int threads, orc_sbvh_stats* stats)
{
    Builder b;
    b.v = v4;
    std::vector<Ref> refs(n);
    Box root;
    for (uint32_t i = 0; i < n; ++i) {
        refs[i].tri = (int)i;
        for (int k = 0; k < 3; ++k) refs[i].box.grow(v4 + ((size_t)i * 3 + k) * 4);
        root.grow(refs[i].box);
    }
    b.min_overlap = root.area() * 1.0e-5f;
    int pd = 0;
    while ((1 << pd) < threads) ++pd;
    b.par_depth = threads > 1 ? pd + 2 : 0;
    Node* tree = b.build(std::move(refs), root, 0);
    SbvhHandle* h = new SbvhHandle();
    int max_depth = 0;
    double s1 = 0.0, s2 = 0.0;
    flatten(tree, h->f, (double)root.area(), &s1, &s2, &max_depth);
    free_tree(tree);
    h->stats.nodes = (uint32_t)h->f.count.size();
    h->stats.tri_refs = (uint32_t)h->f.tri.size();
    h->stats.duplicates = (uint32_t)b.duplicates.load();
    h->stats.depth = max_depth;
    h->stats.sah = (float)s1;
    h->stats.sah_bvhrt = (float)s2;
    if (stats) *stats = h->stats;
    return h;
}

extern "C" void orc_sbvh_free(void* handle) { delete (SbvhHandle*)handle; }

extern "C" void orc_sbvh_trace(const void* handle, const float* v4, const float* rays, uint32_t r,
                               int threads, int32_t* results)
{
    const Flat& f = ((const SbvhHandle*)handle)->f;
    // BVH::traceRecursive (BVH.cpp:88-163) with an explicit stack: the root is
    // entered untested; both children's slabs are tested at their parent and
    // the one with the smaller entry is visited first; a leaf accepts a hit
    // with tmin < t < current tmax.
    auto work = [&](uint32_t a, uint32_t b) {
        std::vector<int> stack;
        stack.reserve(128);
        for (uint32_t i = a; i < b; ++i) {
            const float* ray = rays + 8ull * i;
            const float o[3] = {ray[0], ray[1], ray[2]};
            const float d[3] = {ray[4], ray[5], ray[6]};
            const float tmin = ray[3];
            float t = ray[7];
            int32_t id = -1;
            float bu = 0.0f, bv = 0.0f;
            stack.clear();
            stack.push_back(0);
            while (!stack.empty()) {
                const int x = stack.back();
                stack.pop_back();
                if (f.kid[2 * x] < 0) {
                    for (int k = 0; k < f.count[x]; ++k) {
                        const int tri = f.tri[f.first[x] + k];
                        float tt, uu, vv;
                        const float* p = v4 + (size_t)tri * 12;
                        if (orc_ray_triangle(p, p + 4, p + 8, o, d, tmin, t, &tt, &uu, &vv)) {
                            t = tt;
                            bu = uu;
                            bv = vv;
                            id = tri;
                        }
                    }
                    continue;
                }
                int c0 = f.kid[2 * x], c1 = f.kid[2 * x + 1];
                float n0, f0, n1, f1;
                ray_box(&f.lo[3 * c0], &f.hi[3 * c0], o, d, &n0, &f0);
                ray_box(&f.lo[3 * c1], &f.hi[3 * c1], o, d, &n1, &f1);
                bool h0 = n0 <= f0 && f0 >= tmin && n0 <= t;
                bool h1 = n1 <= f1 && f1 >= tmin && n1 <= t;
                if (h0 && h1 && n0 > n1) {
                    std::swap(c0, c1);
                    std::swap(h0, h1);
                }
                if (h1) stack.push_back(c1);
                if (h0) stack.push_back(c0);
            }
            int32_t* out = results + 4ull * i;
            out[0] = id;
            std::memcpy(&out[1], &t, 4);
            std::memcpy(&out[2], &bu, 4);
            std::memcpy(&out[3], &bv, 4);
        }
    };
    if (threads <= 1) {
        work(0, r);
        return;
    }
    std::vector<std::thread> pool;
    const uint32_t chunk = (r + (uint32_t)threads - 1) / (uint32_t)threads;
    for (int k = 0; k < threads; ++k) {
        const uint32_t a = std::min(r, (uint32_t)k * chunk), b = std::min(r, a + chunk);
        pool.emplace_back(work, a, b);
    }
    for (auto& th : pool) th.join();
}
