// Host implementation of the C ABI in include/atrbvh.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "internal.hpp"

namespace atr {

namespace {
thread_local std::string g_last_error;
constexpr int kPinnedSlots = 4096;
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }

atr_status fail(atr_status code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

// Called only from a catch handler: rethrows the exception being handled to
// name it.  Storing the message may itself fail (no memory): the status is
// returned all the same.
atr_status fail_from_exception() noexcept
{
    atr_status code = ATR_ERR_INTERNAL;
    try {
        try {
            throw;
        } catch (const std::bad_alloc&) {
            code = ATR_ERR_OUT_OF_MEMORY;
            g_last_error = "host memory allocation failed";
        } catch (const std::exception& e) {
            g_last_error = std::string("host exception: ") + e.what();
        } catch (...) {
            g_last_error = "unknown host exception";
        }
    } catch (...) {
    }
    return code;
}

// ------------------------------------------------------------- allocator
atr_status dalloc(atr_context* ctx, size_t bytes, void** p)
{
    if (bytes == 0) bytes = 256;
    bytes = (bytes + 255) & ~size_t(255);
    auto it = ctx->free_blocks.find(bytes);
    if (it != ctx->free_blocks.end()) {
        *p = it->second;
        ctx->free_blocks.erase(it);
        ctx->live[*p] = bytes;
        return ATR_OK;
    }
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) {
        // release cached blocks and retry once
        (void)hipGetLastError();
        ATR_HIP(hipStreamSynchronize(ctx->stream));
        for (auto& kv : ctx->free_blocks) (void)hipFree(kv.second);
        ctx->free_blocks.clear();
        e = hipMalloc(&q, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(ATR_ERR_OUT_OF_MEMORY,
                        "device allocation of " + std::to_string(bytes) + " bytes failed");
        }
    }
    ctx->live[q] = bytes;
    *p = q;
    return ATR_OK;
}

// Returns a block to the cache.  Work on the context stream that still uses
// it is ordered before any later user of the same block.
void dfree(atr_context* ctx, void* p)
{
    if (!p) return;
    auto it = ctx->live.find(p);
    if (it == ctx->live.end()) return;
    ctx->free_blocks.emplace(it->second, p);
    ctx->live.erase(it);
}

// ------------------------------------------------------------- profiling
static hipEvent_t take_event(atr_context* ctx)
{
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

LaunchScope::LaunchScope(atr_context* c, const char* name, double bytes) : ctx(c), rec(-1)
{
    if (!ctx->profiling || ctx->capturing || ctx->no_alloc) return;
    auto it = ctx->prof_index.find(name);
    if (it == ctx->prof_index.end()) {
        ctx->prof.emplace_back();
        ctx->prof.back().name = name;
        rec = (int)ctx->prof.size() - 1;
        ctx->prof_index[name] = rec;
    } else {
        rec = it->second;
    }
    ctx->prof[rec].bytes += bytes;
    ctx->prof[rec].launches += 1;
    e0 = take_event(ctx);
    e1 = take_event(ctx);
    if (e0) (void)hipEventRecord(e0, ctx->stream);
}

atr_status LaunchScope::finish()
{
    hipError_t e = hipGetLastError();
    if (rec >= 0 && e0 && e1) {
        (void)hipEventRecord(e1, ctx->stream);
        ctx->prof[rec].pending.emplace_back(e0, e1);
    }
    if (e != hipSuccess)
        return fail(ATR_ERR_DEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
    return ATR_OK;
}

static void resolve_profile(atr_context* ctx)
{
    for (auto& r : ctx->prof) {
        for (auto& pr : r.pending) {
            float ms = 0.0f;
            (void)hipEventSynchronize(pr.second);
            if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) r.total_ms += ms;
            ctx->event_pool.push_back(pr.first);
            ctx->event_pool.push_back(pr.second);
        }
        r.pending.clear();
    }
}

}  // namespace atr

using namespace atr;

// ================================================================== C ABI
extern "C" {

const char* atr_last_error(void) { return g_last_error.c_str(); }

const char* atr_version(void) { return "atrbvh-mi355x 0.1 (gfx950)"; }

atr_status atr_context_create(int device, atr_context** out)
try {
    if (!out) return fail(ATR_ERR_INVALID_ARGUMENT, "out is NULL");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return fail(ATR_ERR_DEVICE, "no HIP device available");
    }
    if (device < 0 || device >= count) return fail(ATR_ERR_INVALID_ARGUMENT, "bad device index");
    ATR_HIP(hipSetDevice(device));
    atr_context* ctx = new atr_context();
    ctx->device = device;
    ctx->spin_limit = kSpinLimit;
    if (const char* lim = std::getenv("ATRBVH_SPIN_LIMIT")) {
        char* end = nullptr;
        const unsigned long v = std::strtoul(lim, &end, 10);
        if (end != lim && *end == '\0' && v <= kSpinLimit) ctx->spin_limit = (uint32_t)v;
    }
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(ATR_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    if (dalloc_t(ctx, 64, &ctx->d_status) != ATR_OK ||
        hipMemsetAsync(ctx->d_status, 0, 256, ctx->stream) != hipSuccess ||
        hipHostMalloc(&ctx->h_pinned, kPinnedSlots * sizeof(CollapseSummary)) != hipSuccess) {
        atr_context_destroy(ctx);
        return fail(ATR_ERR_DEVICE, "context setup failed");
    }
    for (int i = kPinnedSlots - 1; i >= 0; --i) ctx->pinned_free.push_back(i);
    // The no-allocation entry points' scratch: a failed lookup only disables
    // them (they report it), the rest of the library does not need it.
    if (noalloc_setup(device, &ctx->noalloc_scratch) != ATR_OK) {
        ctx->noalloc_error = g_last_error;
        ctx->noalloc_scratch = nullptr;
    }
    *out = ctx;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_context_destroy(atr_context* ctx)
try {
    if (!ctx) return ATR_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    resolve_profile(ctx);
    for (auto& kv : ctx->live) (void)hipFree(kv.first);
    for (auto& kv : ctx->free_blocks) (void)hipFree(kv.second);
    for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->h_pinned) (void)hipHostFree(ctx->h_pinned);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_sync(atr_context* ctx)
try {
    if (!ctx) return fail(ATR_ERR_INVALID_ARGUMENT, "ctx is NULL");
    ATR_HIP(hipStreamSynchronize(ctx->stream));
    uint32_t st = 0;
    ATR_HIP(hipMemcpy(&st, ctx->d_status, 4, hipMemcpyDeviceToHost));
    if (st == 0u) return ATR_OK;
    // Reported once: the word is cleared so that later work on this context
    // is judged on its own.
    ATR_HIP(hipMemsetAsync(ctx->d_status, 0, 4, ctx->stream));
    ATR_HIP(hipStreamSynchronize(ctx->stream));
    if (st & 1u)
        return fail(ATR_ERR_INTERNAL,
                    "a bounded device wait gave up (the sort's results were discarded and "
                    "trees built since the last synchronisation use input order)");
    if (st & 4u) return fail(ATR_ERR_INTERNAL, "inconsistent traversal layout counts");
    return fail(ATR_ERR_UNSUPPORTED, "a tree was deeper than the traversal stack");
}
ATR_CATCH_ALL

void* atr_context_stream(atr_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

atr_status atr_malloc(atr_context* ctx, size_t bytes, void** dptr)
try {
    if (!ctx || !dptr) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    return dalloc(ctx, bytes, dptr);
}
ATR_CATCH_ALL

atr_status atr_free(atr_context* ctx, void* dptr)
try {
    if (!ctx) return fail(ATR_ERR_INVALID_ARGUMENT, "ctx is NULL");
    dfree(ctx, dptr);
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_memcpy_h2d(atr_context* ctx, void* dst, const void* src, size_t bytes)
try {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL");
    if (bytes == 0) return ATR_OK;
    ATR_HIP(hipStreamSynchronize(ctx->stream));
    ATR_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_memcpy_d2h(atr_context* ctx, void* dst, const void* src, size_t bytes)
try {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL");
    if (bytes == 0) return ATR_OK;
    ATR_HIP(hipStreamSynchronize(ctx->stream));
    ATR_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_memset_d32(atr_context* ctx, void* dst, uint32_t value, size_t count)
try {
    if (!ctx || (!dst && count)) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL");
    if (count == 0) return ATR_OK;
    ATR_HIP(hipMemsetD32Async((hipDeviceptr_t)dst, (int)value, count, ctx->stream));
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_profile_enable(atr_context* ctx, int on)
try {
    if (!ctx) return fail(ATR_ERR_INVALID_ARGUMENT, "ctx is NULL");
    ctx->profiling = on != 0;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_profile_reset(atr_context* ctx)
try {
    if (!ctx) return fail(ATR_ERR_INVALID_ARGUMENT, "ctx is NULL");
    resolve_profile(ctx);
    ctx->prof.clear();
    ctx->prof_index.clear();
    return ATR_OK;
}
ATR_CATCH_ALL

int atr_profile_count(atr_context* ctx) { return ctx ? (int)ctx->prof.size() : 0; }

atr_status atr_profile_get(atr_context* ctx, int i, const char** name, int* launches,
                           double* total_ms, double* bytes)
try {
    if (!ctx || i < 0 || i >= (int)ctx->prof.size())
        return fail(ATR_ERR_INVALID_ARGUMENT, "bad profile index");
    resolve_profile(ctx);
    const ProfileRecord& r = ctx->prof[i];
    if (name) *name = r.name.c_str();
    if (launches) *launches = r.launches;
    if (total_ms) *total_ms = r.total_ms;
    if (bytes) *bytes = r.bytes;
    return ATR_OK;
}
ATR_CATCH_ALL

// ------------------------------------------------------------------ scene
static atr_status scene_finish(atr_context* ctx, atr_scene* s, const float* bbmin,
                               const float* bbmax)
{
    ATR_TRY(dalloc_t(ctx, 8, &s->d_bbox));
    if (bbmin && bbmax) {
        for (int a = 0; a < 3; ++a) {
            s->h_bbox[a] = bbmin[a];
            s->h_bbox[3 + a] = bbmax[a];
        }
        s->bbox_on_host = true;
        ATR_HIP(hipMemcpyAsync(s->d_bbox, s->h_bbox, 6 * sizeof(float), hipMemcpyHostToDevice,
                               ctx->stream));
        ATR_HIP(hipStreamSynchronize(ctx->stream));
    } else {
        ATR_TRY(run_scene_bbox(ctx, s->verts, s->n, s->d_bbox));
    }
    return ATR_OK;
}

atr_status atr_scene_create(atr_context* ctx, const float* vertices, uint32_t n,
                            const float* bbmin, const float* bbmax, atr_scene** out)
try {
    // trees index their 2n - 1 nodes with int32
    if (n >= (1u << 30)) return fail(ATR_ERR_UNSUPPORTED, "at most 2^30-1 triangles per scene");
    if (!ctx || !vertices || !out) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n < 2) return fail(ATR_ERR_INVALID_ARGUMENT, "a scene needs at least 2 triangles");
    if ((bbmin == nullptr) != (bbmax == nullptr))
        return fail(ATR_ERR_INVALID_ARGUMENT, "bbmin and bbmax must both be given or both NULL");
    for (size_t i = 0; i < 3ull * n; ++i)
        for (int a = 0; a < 3; ++a)
            if (!std::isfinite(vertices[i * 4 + a]))
                return fail(ATR_ERR_INVALID_ARGUMENT,
                            "non-finite vertex coordinate at vertex " + std::to_string(i));
    atr_scene* s = new atr_scene();
    s->ctx = ctx;
    s->n = n;
    atr_status st = dalloc_t(ctx, 3ull * n, &s->verts);
    if (st != ATR_OK) {
        delete s;
        return st;
    }
    hipError_t e = hipMemcpyAsync(s->verts, vertices, 3ull * n * sizeof(float4),
                                  hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        atr_scene_destroy(s);
        return fail(ATR_ERR_DEVICE, std::string("vertex upload: ") + hipGetErrorString(e));
    }
    st = scene_finish(ctx, s, bbmin, bbmax);
    if (st != ATR_OK) {
        atr_scene_destroy(s);
        return st;
    }
    *out = s;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_scene_create_device(atr_context* ctx, const float* d_vertices, uint32_t n,
                                   const float* bbmin, const float* bbmax, atr_scene** out)
try {
    // trees index their 2n - 1 nodes with int32
    if (n >= (1u << 30)) return fail(ATR_ERR_UNSUPPORTED, "at most 2^30-1 triangles per scene");
    if (!ctx || !d_vertices || !out) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n < 2) return fail(ATR_ERR_INVALID_ARGUMENT, "a scene needs at least 2 triangles");
    if ((bbmin == nullptr) != (bbmax == nullptr))
        return fail(ATR_ERR_INVALID_ARGUMENT, "bbmin and bbmax must both be given or both NULL");
    atr_scene* s = new atr_scene();
    s->ctx = ctx;
    s->n = n;
    atr_status st = dalloc_t(ctx, 3ull * n, &s->verts);
    if (st != ATR_OK) {
        delete s;
        return st;
    }
    hipError_t e = hipMemcpyAsync(s->verts, d_vertices, 3ull * n * sizeof(float4),
                                  hipMemcpyDeviceToDevice, ctx->stream);
    if (e != hipSuccess) {
        atr_scene_destroy(s);
        return fail(ATR_ERR_DEVICE, std::string("vertex copy: ") + hipGetErrorString(e));
    }
    st = scene_finish(ctx, s, bbmin, bbmax);
    if (st != ATR_OK) {
        atr_scene_destroy(s);
        return st;
    }
    *out = s;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_scene_bounds(atr_scene* s, float bbmin[3], float bbmax[3])
try {
    if (!s || !bbmin || !bbmax) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!s->bbox_on_host) {
        ATR_HIP(hipStreamSynchronize(s->ctx->stream));
        ATR_HIP(hipMemcpy(s->h_bbox, s->d_bbox, 6 * sizeof(float), hipMemcpyDeviceToHost));
        s->bbox_on_host = true;
    }
    for (int a = 0; a < 3; ++a) {
        bbmin[a] = s->h_bbox[a];
        bbmax[a] = s->h_bbox[3 + a];
    }
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_scene_vertices(atr_scene* s, float** d_vertices)
try {
    if (!s || !d_vertices) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    *d_vertices = (float*)s->verts;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_scene_destroy(atr_scene* s)
try {
    if (!s) return ATR_OK;
    dfree(s->ctx, s->verts);
    dfree(s->ctx, s->d_bbox);
    delete s;
    return ATR_OK;
}
ATR_CATCH_ALL

// ------------------------------------------------------------------ trees
static atr_status tree_alloc(atr_context* ctx, uint32_t n, atr_tree** out)
{
    atr_tree* t = new atr_tree();
    t->ctx = ctx;
    t->n = n;
    t->first_leaf = n - 1;
    const size_t m = 2ull * n - 1;
    atr_status s = ATR_OK;
    if (s == ATR_OK) s = dalloc_t(ctx, m, &t->parent);
    if (s == ATR_OK) s = dalloc_t(ctx, m, &t->left);
    if (s == ATR_OK) s = dalloc_t(ctx, m, &t->right);
    if (s == ATR_OK) s = dalloc_t(ctx, m, &t->data);
    if (s == ATR_OK) s = dalloc_t(ctx, m, &t->bbmin);
    if (s == ATR_OK) s = dalloc_t(ctx, m, &t->bbmax);
    if (s == ATR_OK) s = dalloc_t(ctx, m, &t->area);
    if (s == ATR_OK) s = dalloc_t(ctx, m, &t->span);
    if (s != ATR_OK) {
        atr_tree_destroy(t);
        return s;
    }
    *out = t;
    return ATR_OK;
}

// Before t's parent / left / right / bbmin / bbmax change: the collapsed tree
// reading them in place gets its own copy (queued on the stream, so it copies
// the values the collapse saw).
static atr_status tree_unshare(atr_tree* t)
{
    atr_collapsed* c = t->sharer;
    if (!c) return ATR_OK;
    atr_context* ctx = t->ctx;
    const size_t m = 2ull * t->n - 1;
    int32_t *p = nullptr, *l = nullptr, *r = nullptr;
    float4 *mn = nullptr, *mx = nullptr;
    atr_status s = dalloc_t(ctx, m, &p);
    if (s == ATR_OK) s = dalloc_t(ctx, m, &l);
    if (s == ATR_OK) s = dalloc_t(ctx, m, &r);
    if (s == ATR_OK) s = dalloc_t(ctx, m, &mn);
    if (s == ATR_OK) s = dalloc_t(ctx, m, &mx);
    hipError_t e = hipSuccess;
    if (s == ATR_OK) e = hipMemcpyAsync(p, t->parent, m * 4, hipMemcpyDeviceToDevice, ctx->stream);
    if (s == ATR_OK && e == hipSuccess)
        e = hipMemcpyAsync(l, t->left, m * 4, hipMemcpyDeviceToDevice, ctx->stream);
    if (s == ATR_OK && e == hipSuccess)
        e = hipMemcpyAsync(r, t->right, m * 4, hipMemcpyDeviceToDevice, ctx->stream);
    if (s == ATR_OK && e == hipSuccess)
        e = hipMemcpyAsync(mn, t->bbmin, m * 16, hipMemcpyDeviceToDevice, ctx->stream);
    if (s == ATR_OK && e == hipSuccess)
        e = hipMemcpyAsync(mx, t->bbmax, m * 16, hipMemcpyDeviceToDevice, ctx->stream);
    if (s == ATR_OK && e != hipSuccess) s = fail(ATR_ERR_DEVICE, "collapsed tree copy failed");
    if (s != ATR_OK) {
        dfree(ctx, p);
        dfree(ctx, l);
        dfree(ctx, r);
        dfree(ctx, mn);
        dfree(ctx, mx);
        return s;
    }
    c->parent = p;
    c->left = l;
    c->right = r;
    c->bbmin = mn;
    c->bbmax = mx;
    c->links_of = nullptr;
    t->sharer = nullptr;
    return ATR_OK;
}

atr_status atr_tree_destroy(atr_tree* t)
try {
    if (!t) return ATR_OK;
    atr_context* ctx = t->ctx;
    if (atr_collapsed* c = t->sharer) {  // the collapsed tree keeps the arrays
        c->links_of = nullptr;
        t->parent = t->left = t->right = nullptr;
        t->bbmin = t->bbmax = nullptr;
    }
    dfree(ctx, t->parent);
    dfree(ctx, t->left);
    dfree(ctx, t->right);
    dfree(ctx, t->data);
    dfree(ctx, t->bbmin);
    dfree(ctx, t->bbmax);
    dfree(ctx, t->area);
    dfree(ctx, t->span);
    delete t;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_lbvh_build(atr_scene* scene, int morton_bits, atr_tree** out)
try {
    if (!scene || !out) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (morton_bits != 30 && morton_bits != 63)
        return fail(ATR_ERR_INVALID_ARGUMENT, "morton_bits must be 30 or 63");
    atr_context* ctx = scene->ctx;
    const uint32_t n = scene->n;
    const int kb = morton_bits == 63 ? 8 : 4;
    const int passes = morton_bits == 63 ? 8 : 4;
    atr_tree* t = nullptr;
    ATR_TRY(tree_alloc(ctx, n, &t));
    void* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    uint32_t* hist = nullptr;
    uint64_t* slots = nullptr;
    const uint32_t blocks = std::max(encode_blocks(n), 1u);
    atr_status s = ATR_OK;
    if (s == ATR_OK) s = dalloc(ctx, (size_t)n * kb, &keys[0]);
    if (s == ATR_OK) s = dalloc(ctx, (size_t)n * kb, &keys[1]);
    if (s == ATR_OK) s = dalloc_t(ctx, n, &vals[0]);
    if (s == ATR_OK) s = dalloc_t(ctx, n, &vals[1]);
    if (s == ATR_OK) s = dalloc_t(ctx, (size_t)blocks * passes * 256, &hist);
    if (s == ATR_OK) s = dalloc_t(ctx, n - 1, &slots);
    if (s == ATR_OK)
        s = run_encode(ctx, scene->verts, n, scene->d_bbox, morton_bits, keys[0], nullptr, hist,
                       blocks, slots, n - 1, nullptr, 0);
    bool in_alt = false;
    if (s == ATR_OK)
        s = run_sort(ctx, kb * 8, passes, keys[0], vals[0], n, keys[1], vals[1], true, hist,
                     blocks, nullptr, &in_alt);
    if (s == ATR_OK)
        s = run_build(ctx, morton_bits, keys[in_alt ? 1 : 0], vals[in_alt ? 1 : 0], n,
                      scene->verts, t, slots);
    dfree(ctx, keys[0]);
    dfree(ctx, keys[1]);
    dfree(ctx, vals[0]);
    dfree(ctx, vals[1]);
    dfree(ctx, hist);
    dfree(ctx, slots);
    if (s != ATR_OK) {
        atr_tree_destroy(t);
        return s;
    }
    t->scene = scene;
    *out = t;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_lbvh_build_boxes(atr_scene* scene, const float* d_bbmin, const float* d_bbmax,
                                const int32_t* d_tri, uint32_t num_boxes, int morton_bits,
                                atr_tree** out)
try {
    if (!scene || !d_bbmin || !d_bbmax || !d_tri || !out)
        return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (morton_bits != 30 && morton_bits != 63)
        return fail(ATR_ERR_INVALID_ARGUMENT, "morton_bits must be 30 or 63");
    if (num_boxes < 2) return fail(ATR_ERR_INVALID_ARGUMENT, "at least 2 boxes are needed");
    if (num_boxes > (1u << 30)) return fail(ATR_ERR_INVALID_ARGUMENT, "too many boxes");
    atr_context* ctx = scene->ctx;
    const uint32_t n = num_boxes;
    const int kb = morton_bits == 63 ? 8 : 4;
    const int passes = morton_bits == 63 ? 8 : 4;
    atr_tree* t = nullptr;
    ATR_TRY(tree_alloc(ctx, n, &t));
    void* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    uint32_t* hist = nullptr;
    uint64_t* slots = nullptr;
    const uint32_t blocks = std::max(encode_blocks(n), 1u);
    atr_status s = ATR_OK;
    if (s == ATR_OK) s = dalloc(ctx, (size_t)n * kb, &keys[0]);
    if (s == ATR_OK) s = dalloc(ctx, (size_t)n * kb, &keys[1]);
    if (s == ATR_OK) s = dalloc_t(ctx, n, &vals[0]);
    if (s == ATR_OK) s = dalloc_t(ctx, n, &vals[1]);
    if (s == ATR_OK) s = dalloc_t(ctx, (size_t)blocks * passes * 256, &hist);
    if (s == ATR_OK) s = dalloc_t(ctx, n - 1, &slots);
    const float4* bmn = (const float4*)d_bbmin;
    const float4* bmx = (const float4*)d_bbmax;
    if (s == ATR_OK)
        s = run_encode_boxes(ctx, bmn, bmx, n, scene->d_bbox, morton_bits, keys[0], hist, blocks,
                             slots, n - 1);
    bool in_alt = false;
    if (s == ATR_OK)
        s = run_sort(ctx, kb * 8, passes, keys[0], vals[0], n, keys[1], vals[1], true, hist,
                     blocks, nullptr, &in_alt);
    if (s == ATR_OK)
        s = run_build_boxes(ctx, morton_bits, keys[in_alt ? 1 : 0], vals[in_alt ? 1 : 0], n, bmn,
                            bmx, d_tri, t, slots);
    dfree(ctx, keys[0]);
    dfree(ctx, keys[1]);
    dfree(ctx, vals[0]);
    dfree(ctx, vals[1]);
    dfree(ctx, hist);
    dfree(ctx, slots);
    if (s != ATR_OK) {
        atr_tree_destroy(t);
        return s;
    }
    t->scene = scene;
    t->box_leaves = true;
    *out = t;
    return ATR_OK;
}
ATR_CATCH_ALL

// The radix tree + boxes stage over caller-sorted keys (either leaf source).
static atr_status build_stage(atr_scene* scene, const void* keys, const uint32_t* idx,
                              uint32_t n, int morton_bits, const float4* bmn, const float4* bmx,
                              const int32_t* tri, atr_tree** out)
{
    if (morton_bits != 30 && morton_bits != 63)
        return fail(ATR_ERR_INVALID_ARGUMENT, "morton_bits must be 30 or 63");
    if (n < 2) return fail(ATR_ERR_INVALID_ARGUMENT, "a tree needs at least 2 leaves");
    if (n > (1u << 30)) return fail(ATR_ERR_INVALID_ARGUMENT, "too many leaves");
    atr_context* ctx = scene->ctx;
    atr_tree* t = nullptr;
    ATR_TRY(tree_alloc(ctx, n, &t));
    uint64_t* slots = nullptr;
    atr_status s = dalloc_t(ctx, n - 1, &slots);
    if (s == ATR_OK) s = clear_slots(ctx, slots, n - 1);
    if (s == ATR_OK)
        s = bmn ? run_build_boxes(ctx, morton_bits, keys, idx, n, bmn, bmx, tri, t, slots)
                : run_build(ctx, morton_bits, keys, idx, n, scene->verts, t, slots);
    dfree(ctx, slots);
    if (s != ATR_OK) {
        atr_tree_destroy(t);
        return s;
    }
    t->scene = scene;
    t->box_leaves = bmn != nullptr;
    *out = t;
    return ATR_OK;
}

atr_status atr_build_tree(atr_scene* scene, const void* d_sorted_keys,
                          const uint32_t* d_sort_indices, uint32_t num_triangles, int morton_bits,
                          atr_tree** out)
try {
    if (!scene || !d_sorted_keys || !d_sort_indices || !out)
        return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (num_triangles != scene->n)
        return fail(ATR_ERR_INVALID_ARGUMENT, "num_triangles differs from the scene's");
    return build_stage(scene, d_sorted_keys, d_sort_indices, num_triangles, morton_bits, nullptr,
                       nullptr, nullptr, out);
}
ATR_CATCH_ALL

atr_status atr_build_tree_boxes(atr_scene* scene, const void* d_sorted_keys,
                                const uint32_t* d_routing, const float* d_bbmin,
                                const float* d_bbmax, const int32_t* d_tri, uint32_t num_boxes,
                                int morton_bits, atr_tree** out)
try {
    if (!scene || !d_sorted_keys || !d_routing || !d_bbmin || !d_bbmax || !d_tri || !out)
        return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    return build_stage(scene, d_sorted_keys, d_routing, num_boxes, morton_bits,
                       (const float4*)d_bbmin, (const float4*)d_bbmax, d_tri, out);
}
ATR_CATCH_ALL

atr_status atr_tree_bind_scene(atr_tree* t, atr_scene* scene)
try {
    if (!t) return fail(ATR_ERR_INVALID_ARGUMENT, "tree is NULL");
    if (t->box_leaves)
        return fail(ATR_ERR_UNSUPPORTED, "a tree built from primitive boxes keeps its scene");
    if (scene && scene->n != t->n) return fail(ATR_ERR_INVALID_ARGUMENT, "size mismatch");
    t->scene = scene;
    return ATR_OK;
}
ATR_CATCH_ALL

// `iterations` passes at gamma, 2 gamma, 4 gamma, ...
static atr_status treelet_optimize(atr_tree* t, bool agg, int treelet, int gamma, int iterations,
                                   float ci, float ct)
{
    atr_context* ctx = t->ctx;
    if (iterations > 0) ATR_TRY(tree_unshare(t));
    const uint32_t n = t->n;
    float* nsah = nullptr;
    uint64_t* slots = nullptr;
    int4* rec = nullptr;
    int32_t* d_sched = nullptr;
    int rounds = 0;
    atr_status s = ATR_OK;
    if (s == ATR_OK) s = dalloc_t(ctx, 2ull * n - 1, &nsah);
    if (s == ATR_OK) s = dalloc_t(ctx, n, &slots);
    if (s == ATR_OK && iterations > 0) {
        s = dalloc_t(ctx, t->first_leaf, &rec);
        if (s == ATR_OK) s = run_make_records(ctx, t, rec);
    }
    if (s == ATR_OK && !agg) {
        std::vector<int32_t> sched;
        rounds = trbvh_schedule(treelet, 32, sched);
        if (rounds <= 0) s = fail(ATR_ERR_INTERNAL, "TRBVH schedule generation failed");
        if (s == ATR_OK) s = dalloc_t(ctx, sched.size(), &d_sched);
        if (s == ATR_OK) {
            hipError_t e = hipMemcpyAsync(d_sched, sched.data(), sched.size() * 4,
                                          hipMemcpyHostToDevice, ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) s = fail(ATR_ERR_DEVICE, "schedule upload failed");
        }
    }
    for (int it = 0; it < iterations && s == ATR_OK; ++it) {
        if (agg)
            s = run_atrbvh_pass(ctx, t, treelet, gamma, ci, ct, nsah, slots, rec);
        else
            s = run_trbvh_pass(ctx, t, gamma, ci, ct, nsah, slots, rec, d_sched, rounds);
        gamma *= 2;
    }
    dfree(ctx, nsah);
    dfree(ctx, slots);
    dfree(ctx, rec);
    dfree(ctx, d_sched);
    return s;
}

atr_status atr_trbvh_optimize(atr_tree* t, int treelet_size, int iterations, float ci, float ct)
try {
    if (!t) return fail(ATR_ERR_INVALID_ARGUMENT, "tree is NULL");
    if (treelet_size != 7)
        return fail(ATR_ERR_UNSUPPORTED, "TRBVH supports treelet size 7 only");
    if (iterations < 0) return fail(ATR_ERR_INVALID_ARGUMENT, "iterations must be >= 0");
    return treelet_optimize(t, false, treelet_size, treelet_size, iterations, ci, ct);
}
ATR_CATCH_ALL

atr_status atr_atrbvh_optimize(atr_tree* t, int treelet_size, int iterations, float ci, float ct)
try {
    if (!t) return fail(ATR_ERR_INVALID_ARGUMENT, "tree is NULL");
    if (treelet_size < 3 || treelet_size > 32)
        return fail(ATR_ERR_INVALID_ARGUMENT, "ATRBVH treelet size must be in [3, 32]");
    if (iterations < 0) return fail(ATR_ERR_INVALID_ARGUMENT, "iterations must be >= 0");
    return treelet_optimize(t, true, treelet_size, treelet_size, iterations, ci, ct);
}
ATR_CATCH_ALL

atr_status atr_trbvh_pass(atr_tree* t, int treelet_size, int gamma, float ci, float ct)
try {
    if (!t) return fail(ATR_ERR_INVALID_ARGUMENT, "tree is NULL");
    if (treelet_size != 7)
        return fail(ATR_ERR_UNSUPPORTED, "TRBVH supports treelet size 7 only");
    if (gamma < treelet_size) return fail(ATR_ERR_INVALID_ARGUMENT, "gamma must be >= treelet_size");
    return treelet_optimize(t, false, treelet_size, gamma, 1, ci, ct);
}
ATR_CATCH_ALL

atr_status atr_atrbvh_pass(atr_tree* t, int treelet_size, int gamma, float ci, float ct)
try {
    if (!t) return fail(ATR_ERR_INVALID_ARGUMENT, "tree is NULL");
    if (treelet_size < 3 || treelet_size > 32)
        return fail(ATR_ERR_INVALID_ARGUMENT, "ATRBVH treelet size must be in [3, 32]");
    if (gamma < treelet_size) return fail(ATR_ERR_INVALID_ARGUMENT, "gamma must be >= treelet_size");
    return treelet_optimize(t, true, treelet_size, gamma, 1, ci, ct);
}
ATR_CATCH_ALL

atr_status atr_tree_num_triangles(const atr_tree* t, uint32_t* n)
try {
    if (!t || !n) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    *n = t->n;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_collapsed_num_triangles(const atr_collapsed* c, uint32_t* n)
try {
    if (!c || !n) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    *n = c->n;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_tree_device_view(atr_tree* t, atr_tree_view* v)
try {
    if (!t || !v) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    ATR_TRY(tree_unshare(t));  // the caller may write through the view
    v->num_triangles = t->n;
    v->root = 0;
    v->parent = t->parent;
    v->left = t->left;
    v->right = t->right;
    v->data = t->data;
    v->bbmin = (float*)t->bbmin;
    v->bbmax = (float*)t->bbmax;
    v->area = t->area;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_tree_download(atr_tree* t, atr_tree_view* h)
try {
    if (!t || !h) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    const size_t m = 2ull * t->n - 1;
    ATR_HIP(hipStreamSynchronize(t->ctx->stream));
    if (h->parent) ATR_HIP(hipMemcpy(h->parent, t->parent, m * 4, hipMemcpyDeviceToHost));
    if (h->left) ATR_HIP(hipMemcpy(h->left, t->left, m * 4, hipMemcpyDeviceToHost));
    if (h->right) ATR_HIP(hipMemcpy(h->right, t->right, m * 4, hipMemcpyDeviceToHost));
    if (h->data) ATR_HIP(hipMemcpy(h->data, t->data, m * 4, hipMemcpyDeviceToHost));
    if (h->bbmin) ATR_HIP(hipMemcpy(h->bbmin, t->bbmin, m * 16, hipMemcpyDeviceToHost));
    if (h->bbmax) ATR_HIP(hipMemcpy(h->bbmax, t->bbmax, m * 16, hipMemcpyDeviceToHost));
    if (h->area) ATR_HIP(hipMemcpy(h->area, t->area, m * 4, hipMemcpyDeviceToHost));
    h->num_triangles = t->n;
    h->root = 0;
    return ATR_OK;
}
ATR_CATCH_ALL

// Every kernel that walks a tree trusts its links, so a tree handed in by a
// caller is checked in full on the host: the BVHRT index convention (root 0,
// internal nodes 0..n-2 with data -1, leaves n-1..2n-2 with no children and
// data a permutation of 0..n-1), every node but the root the child of exactly
// one internal node, parent[] equal to those links, and every node reachable
// from the root (which rules out cycles among internal nodes).
atr_status atr_tree_view_check(const atr_tree_view* h)
try {
    if (!h || !h->parent || !h->left || !h->right || !h->data)
        return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    const uint32_t n = h->num_triangles;
    if (n < 2) return fail(ATR_ERR_INVALID_ARGUMENT, "a tree needs at least 2 triangles");
    if (n >= (1u << 30)) return fail(ATR_ERR_UNSUPPORTED, "at most 2^30-1 triangles per tree");
    if (h->root != 0) return fail(ATR_ERR_INVALID_ARGUMENT, "the root must be node 0");
    const size_t m = 2ull * n - 1;
    std::vector<uint8_t> seen(m, 0);
    for (uint32_t p = 0; p + 1 < n; ++p) {
        if (h->data[p] != -1)
            return fail(ATR_ERR_INVALID_ARGUMENT,
                        "internal node " + std::to_string(p) + " has data != -1");
        for (int32_t ch : {h->left[p], h->right[p]}) {
            if (ch <= 0 || (size_t)ch >= m || seen[ch])
                return fail(ATR_ERR_INVALID_ARGUMENT,
                            "tree view is not a binary tree (node " + std::to_string(p) + ")");
            seen[ch] = 1;
            if (h->parent[ch] != (int32_t)p)
                return fail(ATR_ERR_INVALID_ARGUMENT,
                            "parent[" + std::to_string(ch) + "] does not match the links");
        }
    }
    if (h->parent[0] != -1) return fail(ATR_ERR_INVALID_ARGUMENT, "parent[root] must be -1");
    std::vector<uint8_t> tri(n, 0);
    for (size_t i = n - 1; i < m; ++i) {
        if (h->left[i] != -1 || h->right[i] != -1)
            return fail(ATR_ERR_INVALID_ARGUMENT, "leaf " + std::to_string(i) + " has children");
        const int32_t d = h->data[i];
        if (d < 0 || (uint32_t)d >= n || tri[d])
            return fail(ATR_ERR_INVALID_ARGUMENT,
                        "leaf data must be a permutation of the triangle indices");
        tri[d] = 1;
    }
    // reachability from the root: each node has one parent, so a walk visits
    // every reachable node once
    std::vector<int32_t> stack{0};
    size_t visited = 0;
    while (!stack.empty()) {
        const int32_t x = stack.back();
        stack.pop_back();
        ++visited;
        if ((uint32_t)x + 1 < n) {
            stack.push_back(h->left[x]);
            stack.push_back(h->right[x]);
        }
    }
    if (visited != m)
        return fail(ATR_ERR_INVALID_ARGUMENT,
                    "tree view has nodes unreachable from the root (a cycle)");
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_tree_upload(atr_tree* t, const atr_tree_view* h)
try {
    if (!t || !h || !h->parent || !h->left || !h->right || !h->data || !h->bbmin || !h->bbmax ||
        !h->area)
        return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (h->num_triangles != t->n) return fail(ATR_ERR_INVALID_ARGUMENT, "size mismatch");
    const size_t m = 2ull * t->n - 1;
    ATR_TRY(atr_tree_view_check(h));
    ATR_TRY(tree_unshare(t));
    ATR_HIP(hipStreamSynchronize(t->ctx->stream));
    ATR_HIP(hipMemcpy(t->parent, h->parent, m * 4, hipMemcpyHostToDevice));
    ATR_HIP(hipMemcpy(t->left, h->left, m * 4, hipMemcpyHostToDevice));
    ATR_HIP(hipMemcpy(t->right, h->right, m * 4, hipMemcpyHostToDevice));
    ATR_HIP(hipMemcpy(t->data, h->data, m * 4, hipMemcpyHostToDevice));
    ATR_HIP(hipMemcpy(t->bbmin, h->bbmin, m * 16, hipMemcpyHostToDevice));
    ATR_HIP(hipMemcpy(t->bbmax, h->bbmax, m * 16, hipMemcpyHostToDevice));
    ATR_HIP(hipMemcpy(t->area, h->area, m * 4, hipMemcpyHostToDevice));
    std::vector<int2> sp(m);
    host_spans(t->n, h->left, h->right, sp.data());
    ATR_HIP(hipMemcpy(t->span, sp.data(), m * sizeof(int2), hipMemcpyHostToDevice));
    return ATR_OK;
}
ATR_CATCH_ALL

// Tree created from a host view (for tests and for callers holding a tree).
atr_status atr_tree_create_from_host(atr_context* ctx, const atr_tree_view* h, atr_tree** out)
try {
    if (!ctx || !h || !out || h->num_triangles < 2)
        return fail(ATR_ERR_INVALID_ARGUMENT, "bad argument");
    atr_tree* t = nullptr;
    ATR_TRY(tree_alloc(ctx, h->num_triangles, &t));
    atr_status s = atr_tree_upload(t, h);
    if (s != ATR_OK) {
        atr_tree_destroy(t);
        return s;
    }
    *out = t;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_tree_sah(atr_tree* t, float ci, float ct, float* out)
try {
    if (!t || !out) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    const size_t m = 2ull * t->n - 1;
    std::vector<float> a(m);
    ATR_HIP(hipStreamSynchronize(t->ctx->stream));
    ATR_HIP(hipMemcpy(a.data(), t->area, m * 4, hipMemcpyDeviceToHost));
    // SoABVHTree::SAH order: root first, other internal nodes, then leaves
    const uint32_t n = t->n;
    const float total = a[0];
    float si = total;
    float sl = 0.0f;
    for (uint32_t i = 1; i < n - 1; ++i) si += a[i];
    for (uint32_t i = n - 1; i < 2 * n - 1; ++i) sl += a[i];
    *out = (ci * si + ct * sl) / total;
    return ATR_OK;
}
ATR_CATCH_ALL

// ------------------------------------------------------------------ collapse
// The collapsed tree of `t`.  Its parent / left / right / bbmin / bbmax are
// those of t, so when t has no other reader of them the collapsed tree reads
// t's arrays in place (t->sharer) instead of a copy.
static atr_status collapsed_alloc(atr_context* ctx, atr_tree* t, atr_collapsed** out)
{
    const uint32_t n = t->n;
    atr_collapsed* c = new atr_collapsed();
    c->ctx = ctx;
    c->n = n;
    const size_t m = 2ull * n - 1;
    atr_status s = ATR_OK;
    if (!t->sharer) {
        c->parent = t->parent;
        c->left = t->left;
        c->right = t->right;
        c->bbmin = t->bbmin;
        c->bbmax = t->bbmax;
        c->links_of = t;
        t->sharer = c;
    } else {
        if (s == ATR_OK) s = dalloc_t(ctx, m, &c->parent);
        if (s == ATR_OK) s = dalloc_t(ctx, m, &c->left);
        if (s == ATR_OK) s = dalloc_t(ctx, m, &c->right);
        if (s == ATR_OK) s = dalloc_t(ctx, m, &c->bbmin);
        if (s == ATR_OK) s = dalloc_t(ctx, m, &c->bbmax);
    }
    if (s == ATR_OK) s = dalloc_t(ctx, m, &c->data);
    if (s == ATR_OK) s = dalloc_t(ctx, m, &c->count);
    if (s == ATR_OK) s = dalloc_t(ctx, n, &c->tri_index);
    if (s == ATR_OK) s = dalloc_t(ctx, 4ull * (n - 1), &c->tnodes);
    if (s == ATR_OK) s = dalloc_t(ctx, 6ull * n, &c->ttris);
    if (s == ATR_OK) s = dalloc_t(ctx, 2ull * n, &c->ttri_id);
    if (s == ATR_OK) s = dalloc_t(ctx, 1, &c->d_summary);
    if (s == ATR_OK) {
        c->summary_event = take_event(ctx);
        if (!c->summary_event) s = fail(ATR_ERR_DEVICE, "event creation failed");
    }
    if (s == ATR_OK) {
        if (ctx->pinned_free.empty()) {
            s = fail(ATR_ERR_OUT_OF_MEMORY, "too many live collapsed trees in this context");
        } else {
            c->pinned_slot = ctx->pinned_free.back();
            ctx->pinned_free.pop_back();
        }
    }
    if (s != ATR_OK) {
        atr_collapsed_destroy(c);
        return s;
    }
    *out = c;
    return ATR_OK;
}

atr_status atr_collapsed_destroy(atr_collapsed* c)
try {
    if (!c) return ATR_OK;
    atr_context* ctx = c->ctx;
    if (c->links_of) {  // the arrays are the tree's
        c->links_of->sharer = nullptr;
        c->parent = c->left = c->right = nullptr;
        c->bbmin = c->bbmax = nullptr;
    }
    dfree(ctx, c->parent);
    dfree(ctx, c->left);
    dfree(ctx, c->right);
    dfree(ctx, c->data);
    dfree(ctx, c->count);
    dfree(ctx, c->bbmin);
    dfree(ctx, c->bbmax);
    dfree(ctx, c->tri_index);
    dfree(ctx, c->tnodes);
    dfree(ctx, c->ttris);
    dfree(ctx, c->ttri_id);
    dfree(ctx, c->cnodes);
    dfree(ctx, c->cwoop);
    dfree(ctx, c->crows_index);
    dfree(ctx, c->d_summary);
    dfree(ctx, c->stage_emit_count);
    // The event and pinned slot go back to their pools without waiting: a
    // later user records / copies after this tree's pending work in stream
    // order and waits on its own event.
    if (c->summary_event) ctx->event_pool.push_back(c->summary_event);
    if (c->pinned_slot >= 0) ctx->pinned_free.push_back(c->pinned_slot);
    delete c;
    return ATR_OK;
}
ATR_CATCH_ALL

static atr_status collapsed_finished(const atr_collapsed* c)
{
    if (c->stage_emit_count)
        return fail(ATR_ERR_INVALID_ARGUMENT,
                    "collapsed tree is unfinished: atr_collapse_lists has not run on it");
    return ATR_OK;
}

static atr_status collapsed_summary(atr_collapsed* c)
{
    ATR_TRY(collapsed_finished(c));
    if (c->summary_ready) return ATR_OK;
    ATR_HIP(hipEventSynchronize(c->summary_event));
    c->h_summary = ((const CollapseSummary*)c->ctx->h_pinned)[c->pinned_slot];
    c->summary_ready = true;
    return ATR_OK;
}

// Allocates the collapsed tree of `t` (bound to t's scene's vertices for its
// traversal layout).
static atr_status collapse_begin(atr_tree* t, atr_collapsed** out)
{
    atr_collapsed* c = nullptr;
    ATR_TRY(collapsed_alloc(t->ctx, t, &c));
    c->has_layout = false;
    c->verts = t->scene ? t->scene->verts : nullptr;
    *out = c;
    return ATR_OK;
}

// Queues the summary's readback after the list stage; *sah (nullable) blocks.
static atr_status collapse_finish(atr_collapsed* c, float* sah)
{
    atr_context* ctx = c->ctx;
    CollapseSummary* dst = (CollapseSummary*)ctx->h_pinned + c->pinned_slot;
    hipError_t e = hipMemcpyAsync(dst, c->d_summary, sizeof(CollapseSummary),
                                  hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(c->summary_event, ctx->stream);
    if (e != hipSuccess) return fail(ATR_ERR_DEVICE, "summary readback failed");
    if (sah) {
        ATR_TRY(collapsed_summary(c));
        *sah = c->h_summary.cost_root / c->h_summary.area_root;
    }
    return ATR_OK;
}

atr_status atr_collapse(atr_tree* t, float ci, float ct, atr_collapsed** out, float* sah)
try {
    if (!t || !out) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    atr_collapsed* c = nullptr;
    ATR_TRY(collapse_begin(t, &c));
    atr_status s = run_collapse(t->ctx, t, c, ci, ct);
    if (s == ATR_OK) s = collapse_finish(c, sah);
    if (s != ATR_OK) {
        atr_collapsed_destroy(c);
        return s;
    }
    *out = c;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_collapse_costs(atr_tree* t, float ci, float ct, atr_collapsed** out)
try {
    if (!t || !out) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    atr_collapsed* c = nullptr;
    ATR_TRY(collapse_begin(t, &c));
    atr_status s = dalloc_t(t->ctx, t->n, &c->stage_emit_count);
    if (s == ATR_OK) s = run_collapse_costs(t->ctx, t, c, ci, ct, c->stage_emit_count);
    if (s != ATR_OK) {
        atr_collapsed_destroy(c);
        return s;
    }
    c->stage_tree = t;
    *out = c;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_collapse_lists(atr_tree* t, atr_collapsed* c, float* sah)
try {
    if (!t || !c) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!c->stage_emit_count)
        return fail(ATR_ERR_INVALID_ARGUMENT, "collapsed tree does not wait for its lists");
    if (c->stage_tree != t)
        return fail(ATR_ERR_INVALID_ARGUMENT,
                    "tree is not the one atr_collapse_costs computed the costs of");
    ATR_TRY(run_collapse_lists(c->ctx, t, c, c->stage_emit_count));
    dfree(c->ctx, c->stage_emit_count);
    c->stage_emit_count = nullptr;
    c->stage_tree = nullptr;
    return collapse_finish(c, sah);
}
ATR_CATCH_ALL

atr_status atr_collapsed_stats(atr_collapsed* c, float* sah, int* height, uint32_t* used)
try {
    if (!c) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    ATR_TRY(collapsed_summary(c));
    if (sah) *sah = c->h_summary.cost_root / c->h_summary.area_root;
    if (height) *height = c->h_summary.height;
    if (used) *used = c->h_summary.used;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_collapsed_download(atr_collapsed* c, atr_collapsed_view* h)
try {
    if (!c || !h) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    ATR_TRY(collapsed_summary(c));
    const size_t m = 2ull * c->n - 1;
    ATR_HIP(hipStreamSynchronize(c->ctx->stream));
    if (h->parent) ATR_HIP(hipMemcpy(h->parent, c->parent, m * 4, hipMemcpyDeviceToHost));
    if (h->left) ATR_HIP(hipMemcpy(h->left, c->left, m * 4, hipMemcpyDeviceToHost));
    if (h->right) ATR_HIP(hipMemcpy(h->right, c->right, m * 4, hipMemcpyDeviceToHost));
    if (h->data) ATR_HIP(hipMemcpy(h->data, c->data, m * 4, hipMemcpyDeviceToHost));
    if (h->triangle_count)
        ATR_HIP(hipMemcpy(h->triangle_count, c->count, m * 4, hipMemcpyDeviceToHost));
    if (h->bbmin) ATR_HIP(hipMemcpy(h->bbmin, c->bbmin, m * 16, hipMemcpyDeviceToHost));
    if (h->bbmax) ATR_HIP(hipMemcpy(h->bbmax, c->bbmax, m * 16, hipMemcpyDeviceToHost));
    if (h->triangle_indices)
        ATR_HIP(hipMemcpy(h->triangle_indices, c->tri_index, (size_t)c->n * 4,
                          hipMemcpyDeviceToHost));
    h->num_triangles = c->n;
    h->root = 0;
    h->used = c->h_summary.used;
    return ATR_OK;
}
ATR_CATCH_ALL

// ------------------------------------------------------------------ trace
atr_status atr_trace(atr_collapsed* c, const float* d_rays, uint32_t num_rays, int any_hit,
                     int32_t* d_results)
try {
    // the ray queue numbers 64-ray blocks in 32 bits
    if (num_rays > (1u << 31)) return fail(ATR_ERR_INVALID_ARGUMENT, "at most 2^31 rays per call");
    if (!c || (num_rays && (!d_rays || !d_results)))
        return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    ATR_TRY(atr_collapsed_layout(c));
    ATR_TRY(collapsed_summary(c));
    return run_trace(c->ctx, c, (const float4*)d_rays, num_rays, any_hit, d_results,
                     c->h_summary.height);
}
ATR_CATCH_ALL

atr_status atr_collapsed_layout(atr_collapsed* c)
try {
    if (!c) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    ATR_TRY(collapsed_finished(c));
    if (c->has_layout) return ATR_OK;
    if (!c->verts)
        return fail(ATR_ERR_INVALID_ARGUMENT,
                    "collapsed tree has no traversal layout: its tree was not bound to a scene");
    ATR_TRY(run_layout(c->ctx, c, c->verts));
    c->has_layout = true;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_collapsed_compact(atr_collapsed* c, int node_offset_size_div,
                                 atr_compact_layout* out)
try {
    if (!c || !out) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    ATR_TRY(collapsed_finished(c));
    if (node_offset_size_div != 1 && node_offset_size_div != 16)
        return fail(ATR_ERR_INVALID_ARGUMENT, "node_offset_size_div must be 1 or 16");
    if (!c->verts)
        return fail(ATR_ERR_INVALID_ARGUMENT,
                    "collapsed tree has no vertices: its tree was not bound to a scene");
    ATR_TRY(run_compact(c->ctx, c, c->verts, node_offset_size_div));
    out->nodes = (const float*)c->cnodes;
    out->num_nodes = c->compact_nodes;
    out->tris = (const float*)c->cwoop;
    out->tri_indices = c->crows_index;
    out->num_rows = c->compact_rows;
    out->node_offset_size_div = c->compact_div;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_trace_compact(atr_context* ctx, int num_rays, int any_hit, const float* rays,
                             int32_t* results, const float* nodesA, const float* nodesB,
                             const float* nodesC, const float* nodesD, const float* trisA,
                             const float* trisB, const float* trisC, const int32_t* triIndices)
try {
    (void)nodesB;
    (void)nodesC;
    (void)nodesD;
    (void)trisB;
    (void)trisC;
    if (!ctx || (num_rays > 0 && (!rays || !results || !nodesA || !trisA || !triIndices)))
        return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (num_rays < 0) return fail(ATR_ERR_INVALID_ARGUMENT, "num_rays must be >= 0");
    return run_trace_compact(ctx, num_rays, any_hit, (const float4*)rays, (int2*)results,
                             (const float4*)nodesA, (const float4*)trisA, triIndices);
}
ATR_CATCH_ALL

atr_status atr_rays_pinhole(atr_context* ctx, const float eye[3], const float basis[9],
                            float tan_half_fov, uint32_t width, uint32_t height, float tmax,
                            float* d_rays)
try {
    // the kernel numbers rays in 32 bits
    const uint64_t count = (uint64_t)width * height;
    if (count > (1ull << 31)) return fail(ATR_ERR_INVALID_ARGUMENT, "at most 2^31 rays per grid");
    if (!ctx || !eye || !basis || !d_rays) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (count == 0) return ATR_OK;
    return run_rays_pinhole(ctx, eye, basis, tan_half_fov, width, height, tmax, (float4*)d_rays);
}
ATR_CATCH_ALL

atr_status atr_rays_diffuse(atr_context* ctx, atr_scene* scene, const float* d_primary,
                            const int32_t* d_primary_results, uint32_t num_rays, uint32_t seed,
                            float tmax, float* d_rays)
try {
    if (!ctx || !scene || !d_primary || !d_primary_results || !d_rays)
        return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (num_rays == 0) return ATR_OK;
    return run_rays_diffuse(ctx, scene->verts, (const float4*)d_primary, d_primary_results,
                            num_rays, seed, tmax, (float4*)d_rays);
}
ATR_CATCH_ALL

atr_status atr_rays_morton_sort(atr_context* ctx, const float* d_rays, uint32_t num_rays,
                                const int32_t* d_slot_to_id_in, float* d_rays_out,
                                int32_t* d_slot_to_id, int32_t* d_id_to_slot)
try {
    if (!ctx || !d_rays || !d_rays_out || !d_slot_to_id)
        return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (d_rays_out == d_rays) return fail(ATR_ERR_INVALID_ARGUMENT, "output overlaps input");
    if (num_rays == 0) return ATR_OK;
    if (num_rays > (1u << 30)) return fail(ATR_ERR_INVALID_ARGUMENT, "too many rays");
    return run_ray_sort(ctx, (const float4*)d_rays, num_rays, d_slot_to_id_in,
                        (float4*)d_rays_out, d_slot_to_id, d_id_to_slot);
}
ATR_CATCH_ALL

atr_status atr_rays_morton_keys(atr_context* ctx, const float* d_rays, uint32_t num_rays,
                                uint64_t* d_keys, float* d_box)
try {
    if (!ctx || !d_rays || !d_keys || !d_box)
        return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (num_rays == 0) return ATR_OK;
    return run_ray_keys(ctx, (const float4*)d_rays, num_rays, d_keys, d_box);
}
ATR_CATCH_ALL

atr_status atr_rays_unsort_results(atr_context* ctx, const int32_t* d_sorted_results,
                                   const int32_t* d_slot_to_id, uint32_t num_rays,
                                   int32_t* d_results)
try {
    if (!ctx || !d_sorted_results || !d_slot_to_id || !d_results)
        return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (d_results == d_sorted_results)
        return fail(ATR_ERR_INVALID_ARGUMENT, "output overlaps input");
    if (num_rays == 0) return ATR_OK;
    return run_unsort_results(ctx, (const int4*)d_sorted_results, d_slot_to_id, num_rays,
                              (int4*)d_results);
}
ATR_CATCH_ALL

// ------------------------------------------------------------------ stages
atr_status atr_morton_encode(atr_context* ctx, const float* d_vertices, uint32_t n,
                             const float bbmin[3], const float bbmax[3], int morton_bits,
                             void* d_keys, uint32_t* d_indices)
try {
    if (n >= (1u << 30)) return fail(ATR_ERR_UNSUPPORTED, "at most 2^30-1 triangles per call");
    if (!ctx || !d_vertices || !bbmin || !bbmax || !d_keys)
        return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (morton_bits != 30 && morton_bits != 63)
        return fail(ATR_ERR_INVALID_ARGUMENT, "morton_bits must be 30 or 63");
    if (n == 0) return ATR_OK;
    float* d_bbox = nullptr;
    ATR_TRY(dalloc_t(ctx, 8, &d_bbox));
    float hb[6] = {bbmin[0], bbmin[1], bbmin[2], bbmax[0], bbmax[1], bbmax[2]};
    ATR_HIP(hipMemcpyAsync(d_bbox, hb, sizeof(hb), hipMemcpyHostToDevice, ctx->stream));
    ATR_HIP(hipStreamSynchronize(ctx->stream));
    atr_status s = run_encode(ctx, (const float4*)d_vertices, n, d_bbox, morton_bits, d_keys,
                              d_indices, nullptr, 0, nullptr, 0, nullptr, 0);
    dfree(ctx, d_bbox);
    return s;
}
ATR_CATCH_ALL

atr_status atr_sort_pairs(atr_context* ctx, int key_bits, void* d_keys, uint32_t* d_values,
                          uint32_t n)
try {
    if (!ctx || (n && !d_keys)) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (key_bits != 32 && key_bits != 64)
        return fail(ATR_ERR_INVALID_ARGUMENT, "key_bits must be 32 or 64");
    if (n >= (1u << 30)) return fail(ATR_ERR_UNSUPPORTED, "at most 2^30-1 keys per sort");
    if (n <= 1) return ATR_OK;
    void* kalt = nullptr;
    uint32_t* valt = nullptr;
    uint32_t* vtmp = nullptr;
    const bool identity = d_values == nullptr;
    bool in_alt = false;
    atr_status s = dalloc(ctx, (size_t)n * key_bits / 8, &kalt);
    if (s == ATR_OK) s = dalloc_t(ctx, n, &valt);
    if (s == ATR_OK && identity) s = dalloc_t(ctx, n, &vtmp);
    if (s == ATR_OK)
        s = run_sort(ctx, key_bits, key_bits / 8, d_keys, identity ? vtmp : d_values, n, kalt,
                     valt, identity, nullptr, 0, nullptr, &in_alt);
    if (s == ATR_OK && in_alt) {
        hipError_t e = hipMemcpyAsync(d_keys, kalt, (size_t)n * key_bits / 8,
                                      hipMemcpyDeviceToDevice, ctx->stream);
        if (e == hipSuccess && !identity)
            e = hipMemcpyAsync(d_values, valt, (size_t)n * 4, hipMemcpyDeviceToDevice,
                               ctx->stream);
        if (e != hipSuccess) s = fail(ATR_ERR_DEVICE, "sort copy-back failed");
    }
    dfree(ctx, kalt);
    dfree(ctx, valt);
    dfree(ctx, vtmp);
    return s;
}
ATR_CATCH_ALL

atr_status atr_refit(atr_tree* t, atr_scene* scene)
try {
    if (!t) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (t->box_leaves)
        return fail(ATR_ERR_UNSUPPORTED,
                    "a tree built from primitive boxes cannot be refit from triangle vertices");
    if (!scene) scene = t->scene;
    if (!scene) return fail(ATR_ERR_INVALID_ARGUMENT, "no scene given or bound to the tree");
    if (scene->n != t->n) return fail(ATR_ERR_INVALID_ARGUMENT, "scene/tree size mismatch");
    ATR_TRY(tree_unshare(t));
    uint32_t* counters = nullptr;
    ATR_TRY(dalloc_t(t->ctx, t->n, &counters));
    atr_status s = run_refit(t->ctx, t, scene->verts, counters);
    dfree(t->ctx, counters);
    return s;
}
ATR_CATCH_ALL

// ------------------------------------------------------- differentiation
atr_status atr_hit_adjoint(atr_context* ctx, atr_scene* scene, const float* d_rays,
                           const int32_t* d_hits, uint32_t num_rays, const float* d_grad_tuv,
                           const uint32_t* d_index, uint32_t num_grad_vertices, float* d_grad)
try {
    if (!ctx || !scene || !d_rays || !d_hits || !d_grad_tuv || !d_grad)
        return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (scene->ctx != ctx) return fail(ATR_ERR_INVALID_ARGUMENT, "scene of another context");
    if (!d_index && (uint64_t)num_grad_vertices < 3ull * scene->n)
        return fail(ATR_ERR_INVALID_ARGUMENT,
                    "without an index the gradient needs 3 slots per triangle");
    if (num_rays == 0) return ATR_OK;
    return run_hit_adjoint(ctx, scene->verts, scene->n, (const float4*)d_rays,
                           (const int4*)d_hits, num_rays, d_grad_tuv, d_index, num_grad_vertices,
                           d_grad);
}
ATR_CATCH_ALL

atr_status atr_depth_loss(atr_context* ctx, const int32_t* d_hits, const int32_t* d_target_hits,
                          uint32_t num_rays, int same_triangle, float* d_grad_tuv,
                          double* d_loss)
try {
    if (!ctx || !d_hits || !d_target_hits || !d_grad_tuv || !d_loss)
        return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (num_rays == 0) return ATR_OK;
    return run_depth_loss(ctx, (const int4*)d_hits, (const int4*)d_target_hits, num_rays,
                          same_triangle ? 1 : 0, d_grad_tuv, d_loss);
}
ATR_CATCH_ALL

atr_status atr_mesh_step(atr_context* ctx, float* d_mesh_vertices, uint32_t num_vertices,
                         float* d_grad, float lr, float max_step, const uint32_t* d_index,
                         atr_scene* scene)
try {
    if (!ctx || !d_mesh_vertices || !d_index || !scene)
        return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (scene->ctx != ctx) return fail(ATR_ERR_INVALID_ARGUMENT, "scene of another context");
    if (num_vertices == 0) return fail(ATR_ERR_INVALID_ARGUMENT, "empty mesh");
    return run_mesh_step(ctx, (float4*)d_mesh_vertices, num_vertices, d_grad, lr, max_step,
                         d_index, scene->n, scene->verts);
}
ATR_CATCH_ALL

atr_status atr_tree_sah_device(atr_tree* t, float ci, float ct, float* d_out)
try {
    if (!t || !d_out) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    return run_sah_parallel(t->ctx, nullptr, 1, t->n, (int)t->first_leaf, t->area, ci, ct, d_out);
}
ATR_CATCH_ALL

// --------------------------------------------------------- batched builds
static atr_status batch_create(atr_context* ctx, const float* vertices, bool on_device,
                               const uint32_t* tri_offsets, uint32_t meshes,
                               const float* bboxes, atr_batch** out)
{
    if (!vertices || !tri_offsets || !out) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (meshes == 0) return fail(ATR_ERR_INVALID_ARGUMENT, "a batch needs at least one mesh");
    if (tri_offsets[0] != 0) return fail(ATR_ERR_INVALID_ARGUMENT, "tri_offsets[0] must be 0");
    // in 64 bits: an offset near 2^32 must not wrap the check round
    for (uint32_t m = 0; m < meshes; ++m)
        if ((uint64_t)tri_offsets[m + 1] < (uint64_t)tri_offsets[m] + 2u)
            return fail(ATR_ERR_INVALID_ARGUMENT,
                        "mesh " + std::to_string(m) + " has fewer than 2 triangles");
    const uint32_t n = tri_offsets[meshes];
    if (n >= (1u << 30)) return fail(ATR_ERR_UNSUPPORTED, "at most 2^30-1 triangles per batch");
    if (!ctx) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!on_device) {
        for (size_t i = 0; i < 3ull * n; ++i)
            for (int a = 0; a < 3; ++a)
                if (!std::isfinite(vertices[i * 4 + a]))
                    return fail(ATR_ERR_INVALID_ARGUMENT,
                                "non-finite vertex coordinate at vertex " + std::to_string(i));
    }
    atr_batch* b = new atr_batch();
    b->ctx = ctx;
    b->meshes = meshes;
    b->n = n;
    b->h_seg.assign(tri_offsets, tri_offsets + meshes + 1);
    b->passes = forest_passes(meshes);
    b->hist_blocks = encode_forest_blocks(n, 8);  // the most blocks any encode shape uses
    atr_status s = ATR_OK;
    if (s == ATR_OK) s = dalloc_t(ctx, meshes + 1, &b->d_seg);
    if (s == ATR_OK) s = dalloc_t(ctx, 3ull * n, &b->verts);
    if (s == ATR_OK) s = dalloc_t(ctx, 6ull * meshes, &b->d_bbox);
    if (s == ATR_OK) s = dalloc_t(ctx, 6ull * meshes, &b->bbox_acc);
    if (s == ATR_OK) s = dalloc_t(ctx, n, &b->keys[0]);
    if (s == ATR_OK) s = dalloc_t(ctx, n, &b->keys[1]);
    if (s == ATR_OK) s = dalloc_t(ctx, n, &b->vals[0]);
    if (s == ATR_OK) s = dalloc_t(ctx, n, &b->vals[1]);
    if (s == ATR_OK) s = dalloc_t(ctx, (size_t)b->hist_blocks * b->passes * 256, &b->hist);
    if (s == ATR_OK) s = dalloc_t(ctx, n, &b->slots);
    if (s == ATR_OK) s = dalloc_t(ctx, (size_t)b->passes * 256, &b->sort_offsets);
    if (s == ATR_OK) s = dalloc_t(ctx, sort_status_words(n, b->passes), &b->sort_status);
    if (s == ATR_OK) s = dalloc_t(ctx, n, &b->pass_slots);
    if (s == ATR_OK) s = dalloc_t(ctx, 2ull * n, &b->nsah);
    if (s == ATR_OK) s = dalloc_t(ctx, n, &b->counters);
    if (s == ATR_OK) s = dalloc_t(ctx, n, &b->rec);
    if (s == ATR_OK) s = dalloc_t(ctx, meshes, &b->d_sah);
    if (s == ATR_OK) {
        atr_tree* f = new atr_tree();
        f->ctx = ctx;
        f->n = n;
        f->first_leaf = n;
        b->forest = f;
        const size_t m = 2ull * n;
        if (s == ATR_OK) s = dalloc_t(ctx, m, &f->parent);
        if (s == ATR_OK) s = dalloc_t(ctx, m, &f->left);
        if (s == ATR_OK) s = dalloc_t(ctx, m, &f->right);
        if (s == ATR_OK) s = dalloc_t(ctx, m, &f->data);
        if (s == ATR_OK) s = dalloc_t(ctx, m, &f->bbmin);
        if (s == ATR_OK) s = dalloc_t(ctx, m, &f->bbmax);
        if (s == ATR_OK) s = dalloc_t(ctx, m, &f->area);
        if (s == ATR_OK) s = dalloc_t(ctx, m, &f->span);
    }
    if (s == ATR_OK) {
        hipError_t e = hipMemcpyAsync(b->d_seg, tri_offsets, (meshes + 1ull) * 4,
                                      hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(b->verts, vertices, 3ull * n * sizeof(float4),
                               on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                               ctx->stream);
        if (e == hipSuccess && bboxes) {
            b->bbox_given = true;
            b->h_bbox.assign(bboxes, bboxes + 6ull * meshes);
            e = hipMemcpyAsync(b->d_bbox, bboxes, 6ull * meshes * 4, hipMemcpyHostToDevice,
                               ctx->stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) s = fail(ATR_ERR_DEVICE, std::string("batch upload: ") +
                                                          hipGetErrorString(e));
    }
    if (s != ATR_OK) {
        atr_batch_destroy(b);
        return s;
    }
    *out = b;
    return ATR_OK;
}

atr_status atr_batch_create(atr_context* ctx, const float* vertices, const uint32_t* tri_offsets,
                            uint32_t num_meshes, const float* bboxes, atr_batch** out)
try {
    return batch_create(ctx, vertices, false, tri_offsets, num_meshes, bboxes, out);
}
ATR_CATCH_ALL

atr_status atr_batch_create_device(atr_context* ctx, const float* d_vertices,
                                   const uint32_t* tri_offsets, uint32_t num_meshes,
                                   const float* bboxes, atr_batch** out)
try {
    return batch_create(ctx, d_vertices, true, tri_offsets, num_meshes, bboxes, out);
}
ATR_CATCH_ALL

atr_status atr_batch_vertices(atr_batch* b, float** d_vertices)
try {
    if (!b || !d_vertices) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    *d_vertices = (float*)b->verts;
    return ATR_OK;
}
ATR_CATCH_ALL

// The whole batched build as one stream-ordered sequence of kernels over
// persistent buffers: no allocation, no memset node, no host synchronisation,
// so it can be captured into a graph.
static atr_status batch_enqueue(atr_batch* b, int treelet, int iterations, float ci, float ct)
{
    atr_context* ctx = b->ctx;
    const uint32_t n = b->n;
    if (!b->bbox_given)
        ATR_TRY(run_bbox_forest(ctx, b->verts, b->d_seg, b->meshes, n, b->bbox_acc, b->d_bbox));
    ATR_TRY(run_encode_forest(ctx, b->verts, b->d_seg, b->meshes, n, b->d_bbox, b->keys[0],
                              b->hist, b->slots, n, b->encode_items));
    SortWork sw{b->sort_offsets, b->sort_status};
    bool in_alt = false;
    ATR_TRY(run_sort(ctx, 64, b->passes, b->keys[0], b->vals[0], n, b->keys[1], b->vals[1], true,
                     b->hist, encode_forest_blocks(n, b->encode_items), &sw, &in_alt,
                     b->sort_items));
    const int k = in_alt ? 1 : 0;
    ATR_TRY(run_build_forest(ctx, b->keys[k], b->vals[k], b->verts, b->forest, b->d_seg,
                             b->meshes, b->slots));
    int gamma = treelet;
    if (iterations > 0) ATR_TRY(run_make_records(ctx, b->forest, b->rec));
    for (int it = 0; it < iterations; ++it) {
        ATR_TRY(run_atrbvh_pass(ctx, b->forest, treelet, gamma, ci, ct, b->nsah, b->pass_slots,
                                b->rec, b->treelet_block));
        gamma *= 2;
    }
    return ATR_OK;
}

static void batch_drop_graph(atr_batch* b)
{
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
}

atr_status atr_batch_build(atr_batch* b, int treelet_size, int iterations, float ci, float ct,
                           int use_graph)
try {
    if (!b) return fail(ATR_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (treelet_size < 3 || treelet_size > 32)
        return fail(ATR_ERR_INVALID_ARGUMENT, "ATRBVH treelet size must be in [3, 32]");
    if (iterations < 0) return fail(ATR_ERR_INVALID_ARGUMENT, "iterations must be >= 0");
    atr_context* ctx = b->ctx;
    if (!use_graph) {
        ATR_TRY(batch_enqueue(b, treelet_size, iterations, ci, ct));
        b->built = true;
        return ATR_OK;
    }
    const bool same = b->graph_exec && b->g_treelet == treelet_size &&
                      b->g_iterations == iterations && b->g_ci == ci && b->g_ct == ct;
    if (!same) {
        batch_drop_graph(b);
        ATR_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        ctx->capturing = true;
        atr_status s = batch_enqueue(b, treelet_size, iterations, ci, ct);
        ctx->capturing = false;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (s != ATR_OK) {
            if (g) (void)hipGraphDestroy(g);
            return s;
        }
        if (e != hipSuccess)
            return fail(ATR_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
        b->graph = g;
        e = hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            batch_drop_graph(b);
            return fail(ATR_ERR_DEVICE, std::string("graph instantiate: ") +
                                            hipGetErrorString(e));
        }
        b->g_treelet = treelet_size;
        b->g_iterations = iterations;
        b->g_ci = ci;
        b->g_ct = ct;
    }
    ATR_HIP(hipGraphLaunch(b->graph_exec, ctx->stream));
    b->built = true;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_batch_set_config(atr_batch* b, const atr_batch_config* cfg)
try {
    if (!b || !cfg) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if ((cfg->encode_items != 8 && cfg->encode_items != 16) ||
        (cfg->sort_items != 8 && cfg->sort_items != 16))
        return fail(ATR_ERR_INVALID_ARGUMENT, "encode_items and sort_items must be 8 or 16");
    const int tb = cfg->treelet_block == 0 ? 128 : cfg->treelet_block;
    if (tb != 64 && tb != 128 && tb != 256)
        return fail(ATR_ERR_INVALID_ARGUMENT, "treelet_block must be 64, 128 or 256 (0 = 128)");
    if (cfg->encode_items != b->encode_items || cfg->sort_items != b->sort_items ||
        tb != b->treelet_block)
        batch_drop_graph(b);
    b->encode_items = cfg->encode_items;
    b->sort_items = cfg->sort_items;
    b->treelet_block = tb;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_batch_get_config(const atr_batch* b, atr_batch_config* cfg)
try {
    if (!b || !cfg) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    cfg->encode_items = b->encode_items;
    cfg->sort_items = b->sort_items;
    cfg->treelet_block = b->treelet_block;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_batch_sah(atr_batch* b, float ci, float ct, float* out)
try {
    if (!b || !out) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!b->built) return fail(ATR_ERR_INVALID_ARGUMENT, "the batch has not been built");
    atr_context* ctx = b->ctx;
    ATR_TRY(run_forest_sah(ctx, b->forest, b->d_seg, b->meshes, ci, ct, b->d_sah));
    ATR_HIP(hipMemcpyAsync(out, b->d_sah, (size_t)b->meshes * 4, hipMemcpyDeviceToHost,
                           ctx->stream));
    ATR_HIP(hipStreamSynchronize(ctx->stream));
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_batch_sah_device(atr_batch* b, float ci, float ct, float* d_out)
try {
    if (!b || !d_out) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!b->built) return fail(ATR_ERR_INVALID_ARGUMENT, "the batch has not been built");
    return run_sah_parallel(b->ctx, b->d_seg, b->meshes, 0, (int)b->forest->first_leaf,
                            b->forest->area, ci, ct, d_out);
}
ATR_CATCH_ALL

atr_status atr_batch_trace(atr_batch* b, const float* d_rays, uint32_t rays_per_mesh,
                           int32_t* d_hits)
try {
    if (!b || !d_rays || !d_hits) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!b->built) return fail(ATR_ERR_INVALID_ARGUMENT, "the batch has not been built");
    const uint64_t r = (uint64_t)rays_per_mesh * b->meshes;
    if (rays_per_mesh == 0) return ATR_OK;
    if (r >= (1ull << 32)) return fail(ATR_ERR_UNSUPPORTED, "too many rays");
    return run_trace_forest(b->ctx, b->forest, b->d_seg, b->verts, (const float4*)d_rays,
                            (uint32_t)r, rays_per_mesh, d_hits);
}
ATR_CATCH_ALL

atr_status atr_batch_hit_adjoint(atr_batch* b, const float* d_rays, const int32_t* d_hits,
                                 uint32_t num_rays, const float* d_grad_tuv, float* d_grad)
try {
    if (!b || !d_rays || !d_hits || !d_grad_tuv || !d_grad)
        return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (num_rays == 0) return ATR_OK;
    return run_hit_adjoint(b->ctx, b->verts, b->n, (const float4*)d_rays, (const int4*)d_hits,
                           num_rays, d_grad_tuv, nullptr, 3u * b->n, d_grad);
}
ATR_CATCH_ALL

atr_status atr_batch_step(atr_batch* b, float* d_grad, float lr, float max_step)
try {
    if (!b || !d_grad) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    return run_mesh_step(b->ctx, b->verts, 3u * b->n, d_grad, lr, max_step, nullptr, 0, nullptr);
}
ATR_CATCH_ALL

atr_status atr_batch_refit(atr_batch* b)
try {
    if (!b) return fail(ATR_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (!b->built) return fail(ATR_ERR_INVALID_ARGUMENT, "the batch has not been built");
    return run_refit(b->ctx, b->forest, b->verts, b->counters);
}
ATR_CATCH_ALL

atr_status atr_batch_rebuild(atr_batch* b, const uint32_t* meshes, uint32_t count,
                             int treelet_size, int iterations, float ci, float ct)
try {
    if (!b || (!meshes && count)) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!b->built) return fail(ATR_ERR_INVALID_ARGUMENT, "the batch has not been built");
    if (count == 0) return ATR_OK;
    if (count > b->meshes)
        return fail(ATR_ERR_INVALID_ARGUMENT, "more mesh indices than the batch has meshes");
    atr_context* ctx = b->ctx;
    std::vector<uint32_t> off(count + 1, 0);
    std::vector<uint8_t> seen(b->meshes, 0);
    for (uint32_t j = 0; j < count; ++j) {
        const uint32_t m = meshes[j];
        if (m >= b->meshes || seen[m])
            return fail(ATR_ERR_INVALID_ARGUMENT, "mesh indices must be distinct and in range");
        seen[m] = 1;
        off[j + 1] = off[j] + (b->h_seg[m + 1] - b->h_seg[m]);
    }
    float4* sv = nullptr;
    uint32_t* d_map = nullptr;
    atr_batch* sb = nullptr;
    atr_status s = dalloc_t(ctx, 3ull * off[count], &sv);
    if (s == ATR_OK) s = dalloc_t(ctx, count, &d_map);
    for (uint32_t j = 0; j < count && s == ATR_OK; ++j) {
        const uint32_t m = meshes[j];
        hipError_t e = hipMemcpyAsync(sv + 3ull * off[j], b->verts + 3ull * b->h_seg[m],
                                      3ull * (off[j + 1] - off[j]) * sizeof(float4),
                                      hipMemcpyDeviceToDevice, ctx->stream);
        if (e != hipSuccess) s = fail(ATR_ERR_DEVICE, "sub-batch vertex copy failed");
    }
    if (s == ATR_OK) {
        hipError_t e = hipMemcpyAsync(d_map, meshes, count * 4ull, hipMemcpyHostToDevice,
                                      ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) s = fail(ATR_ERR_DEVICE, "sub-batch map upload failed");
    }
    // A batch built on caller boxes keeps them: the rebuilt meshes quantise
    // their Morton codes in the same frames as a full build does.
    std::vector<float> boxes;
    if (b->bbox_given) {
        boxes.reserve(6ull * count);
        for (uint32_t j = 0; j < count; ++j)
            boxes.insert(boxes.end(), b->h_bbox.begin() + 6ull * meshes[j],
                         b->h_bbox.begin() + 6ull * meshes[j] + 6);
    }
    if (s == ATR_OK)
        s = batch_create(ctx, (const float*)sv, true, off.data(), count,
                         b->bbox_given ? boxes.data() : nullptr, &sb);
    if (s == ATR_OK) s = atr_batch_build(sb, treelet_size, iterations, ci, ct, 0);
    if (s == ATR_OK)
        s = run_forest_splice(ctx, sb->forest, sb->d_seg, count, d_map, b->d_seg, b->forest);
    if (sb) atr_batch_destroy(sb);
    dfree(ctx, sv);
    dfree(ctx, d_map);
    return s;
}
ATR_CATCH_ALL

atr_status atr_batch_extract(atr_batch* b, uint32_t mesh, atr_tree** out)
try {
    if (!b || !out) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (mesh >= b->meshes) return fail(ATR_ERR_INVALID_ARGUMENT, "mesh index out of range");
    if (!b->built) return fail(ATR_ERR_INVALID_ARGUMENT, "the batch has not been built");
    const uint32_t s0 = b->h_seg[mesh];
    const uint32_t nm = b->h_seg[mesh + 1] - s0;
    atr_tree* t = nullptr;
    ATR_TRY(tree_alloc(b->ctx, nm, &t));
    atr_status s = run_forest_extract(b->ctx, b->forest, s0, nm, t);
    if (s != ATR_OK) {
        atr_tree_destroy(t);
        return s;
    }
    *out = t;
    return ATR_OK;
}
ATR_CATCH_ALL

atr_status atr_batch_destroy(atr_batch* b)
try {
    if (!b) return ATR_OK;
    atr_context* ctx = b->ctx;
    (void)hipStreamSynchronize(ctx->stream);
    batch_drop_graph(b);
    if (b->forest) atr_tree_destroy(b->forest);
    dfree(ctx, b->d_seg);
    dfree(ctx, b->verts);
    dfree(ctx, b->d_bbox);
    dfree(ctx, b->bbox_acc);
    dfree(ctx, b->keys[0]);
    dfree(ctx, b->keys[1]);
    dfree(ctx, b->vals[0]);
    dfree(ctx, b->vals[1]);
    dfree(ctx, b->hist);
    dfree(ctx, b->slots);
    dfree(ctx, b->sort_offsets);
    dfree(ctx, b->sort_status);
    dfree(ctx, b->pass_slots);
    dfree(ctx, b->nsah);
    dfree(ctx, b->counters);
    dfree(ctx, b->rec);
    dfree(ctx, b->d_sah);
    delete b;
    return ATR_OK;
}
ATR_CATCH_ALL

}  // extern "C"
