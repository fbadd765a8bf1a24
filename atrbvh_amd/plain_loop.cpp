// Batched build against the plain per-mesh loop, both through the C ABI from
// one C++ host thread (no interpreter overhead on either side).
//
// usage: plain_loop <meshes> <size_seed> <lo> <hi> <soup_seed_base> <reps>
//                   [encode_items sort_items treelet_block]
// The workload is the batch of atr_gen_batch_sizes(size_seed, meshes, lo, hi)
// soups atr_gen_soup_batch(soup_seed_base, ...), uploaded once.  Timed, after
// one untimed pass each:
//   batch: atr_batch_build(treelet 9, 2 iterations, graph replay), reps times;
//   loop:  for every mesh, atr_lbvh_build (30-bit keys) + atr_atrbvh_optimize
//          (9, 2) + atr_tree_destroy on scenes created beforehand, reps times.
// Then every mesh's loop tree SAH is compared with the batch's, bit for bit.
// Prints one JSON line; exit status 0, or 3 on a library error.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../include/atrbvh.h"
#include "../include/atrbvh_scenes.h"

static int die(const char* what, atr_status s)
{
    std::printf("{\"error\": \"%s: status %d: %s\"}\n", what, (int)s, atr_last_error());
    return 3;
}

#define CHECK(expr)                           \
    do {                                      \
        atr_status _s = (expr);               \
        if (_s != ATR_OK) return die(#expr, _s); \
    } while (0)

int main(int argc, char** argv)
{
    if (argc != 7 && argc != 10) return 2;
    const uint32_t meshes = (uint32_t)std::strtoul(argv[1], nullptr, 10);
    const uint64_t size_seed = std::strtoull(argv[2], nullptr, 10);
    const uint32_t lo = (uint32_t)std::strtoul(argv[3], nullptr, 10);
    const uint32_t hi = (uint32_t)std::strtoul(argv[4], nullptr, 10);
    const uint64_t soup_seed = std::strtoull(argv[5], nullptr, 10);
    const int reps = std::atoi(argv[6]);
    atr_batch_config cfg{16, 16, 128};
    if (argc == 10) {
        cfg.encode_items = std::atoi(argv[7]);
        cfg.sort_items = std::atoi(argv[8]);
        cfg.treelet_block = std::atoi(argv[9]);
    }
    if (meshes == 0 || reps < 1) return 2;
    std::vector<uint32_t> off(meshes + 1);
    atr_gen_batch_sizes(size_seed, meshes, lo, hi, off.data());
    const uint32_t n = off[meshes];
    std::vector<float> v(12ull * n);
    atr_gen_soup_batch(soup_seed, meshes, off.data(), 0.005f, v.data());

    atr_context* ctx = nullptr;
    CHECK(atr_context_create(0, &ctx));
    float* d_v = nullptr;
    CHECK(atr_malloc(ctx, v.size() * sizeof(float), (void**)&d_v));
    CHECK(atr_memcpy_h2d(ctx, d_v, v.data(), v.size() * sizeof(float)));
    std::vector<float>().swap(v);

    atr_batch* b = nullptr;
    CHECK(atr_batch_create_device(ctx, d_v, off.data(), meshes, nullptr, &b));
    CHECK(atr_batch_set_config(b, &cfg));
    std::vector<atr_scene*> scenes(meshes, nullptr);
    for (uint32_t m = 0; m < meshes; ++m)
        CHECK(atr_scene_create_device(ctx, d_v + 12ull * off[m], off[m + 1] - off[m], nullptr,
                                      nullptr, &scenes[m]));
    CHECK(atr_sync(ctx));

    using clk = std::chrono::steady_clock;
    // batch: capture + first replay untimed
    CHECK(atr_batch_build(b, 9, 2, 1.2f, 1.0f, 1));
    CHECK(atr_sync(ctx));
    auto t0 = clk::now();
    for (int r = 0; r < reps; ++r) CHECK(atr_batch_build(b, 9, 2, 1.2f, 1.0f, 1));
    CHECK(atr_sync(ctx));
    const double batch_ms =
        std::chrono::duration<double, std::milli>(clk::now() - t0).count() / reps;
    std::vector<float> sah_batch(meshes);
    CHECK(atr_batch_sah(b, 1.2f, 1.0f, sah_batch.data()));

    // loop: one untimed pass fills the allocator cache with every mesh size
    auto loop_once = [&]() -> atr_status {
        for (uint32_t m = 0; m < meshes; ++m) {
            atr_tree* t = nullptr;
            atr_status s = atr_lbvh_build(scenes[m], 30, &t);
            if (s == ATR_OK) s = atr_atrbvh_optimize(t, 9, 2, 1.2f, 1.0f);
            if (t) atr_tree_destroy(t);
            if (s != ATR_OK) return s;
        }
        return ATR_OK;
    };
    CHECK(loop_once());
    CHECK(atr_sync(ctx));
    t0 = clk::now();
    for (int r = 0; r < reps; ++r) CHECK(loop_once());
    CHECK(atr_sync(ctx));
    const double loop_ms =
        std::chrono::duration<double, std::milli>(clk::now() - t0).count() / reps;

    bool same = true;
    for (uint32_t m = 0; m < meshes; ++m) {
        atr_tree* t = nullptr;
        CHECK(atr_lbvh_build(scenes[m], 30, &t));
        CHECK(atr_atrbvh_optimize(t, 9, 2, 1.2f, 1.0f));
        float s = 0.0f;
        CHECK(atr_tree_sah(t, 1.2f, 1.0f, &s));
        same &= std::memcmp(&s, &sah_batch[m], 4) == 0;
        atr_tree_destroy(t);
    }
    std::printf("{\"meshes\": %u, \"triangles\": %u, \"reps\": %d, \"encode_items\": %d, "
                "\"sort_items\": %d, \"treelet_block\": %d, \"batch_ms\": %.4f, "
                "\"loop_ms\": %.4f, "
                "\"speedup\": %.3f, \"sah_identical\": %s}\n",
                meshes, n, reps, cfg.encode_items, cfg.sort_items, cfg.treelet_block, batch_ms,
                loop_ms,
                loop_ms / batch_ms, same ? "true" : "false");
    for (atr_scene* s : scenes) atr_scene_destroy(s);
    atr_batch_destroy(b);
    atr_free(ctx, d_v);
    atr_context_destroy(ctx);
    return 0;
}
