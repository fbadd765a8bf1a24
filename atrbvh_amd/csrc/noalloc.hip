// Entry points that allocate nothing, for callers that cannot call hipMalloc:
// code under HIP graph capture, tight serving loops, harnesses that forbid
// device allocation.  Their scratch memory is one file-scope device array of
// fixed size, of which every GPU the library runs on holds its own copy for
// the life of the process.
//
// The array's name is only a host-side stand-in for the symbol: host code
// uses the device address that hipGetSymbolAddress returns for it, looked up
// once per device (with that device current) when a context is created.
//
// Every call on one GPU shares that GPU's array and the array keeps the last
// call's contents, so everything that must start at zero (the sort's
// look-back status words and digit offsets) is cleared by a kernel on the
// caller's stream at the start of each call.  Calls that overlap on one GPU
// (two contexts, or two host threads, on the same device) race on the array.
#include <mutex>

#include "device.hpp"
#include "internal.hpp"

namespace atr {

namespace {

// Largest sort the array holds: 1,048,576 (key, value) pairs of 64-bit keys.
constexpr uint32_t kNoAllocMaxKeys = 1u << 20;

constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Carving of the array for a sort of n keys of kb bytes (all 256-byte
// aligned): alternate keys, alternate values, identity values, block
// histograms, digit offsets, look-back status words.
struct SortCarve {
    size_t kalt, valt, vtmp, hist, offsets, status, total;
};

SortCarve carve(uint32_t n, int kb)
{
    const int passes = kb;
    SortCarve c{};
    size_t at = 0;
    c.kalt = at;
    at += align256((size_t)n * kb);
    c.valt = at;
    at += align256((size_t)n * 4);
    c.vtmp = at;
    at += align256((size_t)n * 4);
    c.hist = at;
    at += align256(sort_hist_words(n, passes) * 4);
    c.offsets = at;
    at += align256((size_t)passes * 256 * 4);
    c.status = at;
    at += align256(sort_status_words(n, passes) * 4);
    c.total = at;
    return c;
}

// 24 MiB: the carving of the largest sort (64-bit keys, kNoAllocMaxKeys
// pairs) needs 22.0 MiB; atr_sort_pairs_noalloc checks it against this size
// before every call.
constexpr size_t kScratchBytes = 24ull << 20;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_mu;
constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};

}  // namespace

atr_status noalloc_setup(int device, void** out)
{
    if (device < 0 || device >= kMaxDevices)
        return fail(ATR_ERR_UNSUPPORTED, "device index beyond the no-allocation table");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(ATR_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                            hipGetErrorString(e));
        g_scratch[device] = p;
    }
    *out = g_scratch[device];
    return ATR_OK;
}

}  // namespace atr

using namespace atr;

extern "C" atr_status atr_sort_pairs_noalloc(atr_context* ctx, int key_bits, void* d_keys,
                                             uint32_t* d_values, uint32_t n)
try {
    if (!ctx || (n && !d_keys)) return fail(ATR_ERR_INVALID_ARGUMENT, "NULL argument");
    if (key_bits != 32 && key_bits != 64)
        return fail(ATR_ERR_INVALID_ARGUMENT, "key_bits must be 32 or 64");
    if (n > kNoAllocMaxKeys)
        return fail(ATR_ERR_INVALID_ARGUMENT,
                    "atr_sort_pairs_noalloc sorts at most 1048576 pairs (" + std::to_string(n) +
                        " given); use atr_sort_pairs");
    if (!ctx->noalloc_scratch)
        return fail(ATR_ERR_DEVICE, "no-allocation scratch unavailable: " + ctx->noalloc_error);
    if (n <= 1) return ATR_OK;
    const int kb = key_bits / 8;
    const SortCarve c = carve(n, kb);
    if (c.total > kScratchBytes)
        return fail(ATR_ERR_INTERNAL, "no-allocation scratch too small for this sort");
    uint8_t* base = (uint8_t*)ctx->noalloc_scratch;
    SortWork work{(uint32_t*)(base + c.offsets), (uint32_t*)(base + c.status),
                  (uint32_t*)(base + c.hist)};
    const bool identity = d_values == nullptr;
    uint32_t* vals = identity ? (uint32_t*)(base + c.vtmp) : d_values;
    bool in_alt = false;
    ctx->no_alloc = true;
    const atr_status s = run_sort(ctx, key_bits, kb, d_keys, vals, n, base + c.kalt,
                                  (uint32_t*)(base + c.valt), identity, nullptr, 0, &work,
                                  &in_alt);
    ctx->no_alloc = false;
    if (s != ATR_OK) return s;
    // 4 or 8 passes: the last pass writes the caller's buffers
    if (in_alt) return fail(ATR_ERR_INTERNAL, "sort ended in the scratch buffers");
    return ATR_OK;
}
ATR_CATCH_ALL
