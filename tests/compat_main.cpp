// Drives the BVHRT C++ facade the way Renderer::getCudaBVH does
// (gpu-ray-traversal/src/rt/cuda/Renderer.cpp:164-216): scene -> LBVH ->
// optimizer -> collapse, then downloads both trees.
// usage: compat_main <vertices.f32> <n> <morton 0|1> <opt: none|trbvh|atrbvh> <out.bin>
// out.bin: float sah_tree, float sah_collapsed, uint32 used, then the tree's
// parent/left/right/data (int32 x (2n-1) each) and the collapsed tree's
// data/triangleCount (int32 x (2n-1)) and triangleIndices (int32 x used).
// Exit status 0 on success, 3 when the library reports an error (message on
// stdout as "status <code>: <text>").
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bvhrt_compat.hpp"

int main(int argc, char** argv)
{
    if (argc != 6) return 2;
    const unsigned n = (unsigned)std::stoul(argv[2]);
    std::vector<BVHRT::Float4> v(3ull * n);
    FILE* f = std::fopen(argv[1], "rb");
    if (!f || std::fread(v.data(), sizeof(BVHRT::Float4), v.size(), f) != v.size()) return 2;
    std::fclose(f);
    try {
        BVHRT::SceneWrapper scene(n, v.data());
        BVHRT::LBVHBuilder builder(std::atoi(argv[3]) != 0);
        BVHRT::BVHTree* tree = builder.BuildTree(&scene);
        const std::string opt = argv[4];
        if (opt == "trbvh") BVHRT::TRBVHOptimizer(7, 3).Optimize(tree);
        if (opt == "atrbvh") BVHRT::AgglomerativeTreeletOptimizer(9, 2).Optimize(tree);
        BVHRT::BVHTreeInstanceManager manager;
        BVHRT::SoABVHTree* host = manager.DeviceToHostTree(tree);
        float sah = 0.0f;
        BVHRT::DeviceCollapsedTree* c = BVHRT::BVHTreeCollapser().Collapse(tree, &sah);
        BVHRT::BVHCollapsedTree* hc = manager.DeviceToHostCollapsedTree(c);
        FILE* o = std::fopen(argv[5], "wb");
        const float sah_tree = host->SAH();
        std::fwrite(&sah_tree, 4, 1, o);
        std::fwrite(&sah, 4, 1, o);
        std::fwrite(&hc->used, 4, 1, o);
        const size_t m = 2ull * n - 1;
        std::fwrite(host->parentIndices.data(), 4, m, o);
        std::fwrite(host->leftIndices.data(), 4, m, o);
        std::fwrite(host->rightIndices.data(), 4, m, o);
        std::fwrite(host->dataIndices.data(), 4, m, o);
        std::fwrite(hc->nodes.dataIndices.data(), 4, m, o);
        std::fwrite(hc->nodes.triangleCount.data(), 4, m, o);
        std::fwrite(hc->triangleIndices.data(), 4, hc->used, o);
        std::fclose(o);
        manager.FreeDeviceCollapsedTree(c);
        manager.FreeDeviceTree(tree);
        delete host;
        delete hc;
    } catch (const BVHRT::Error& e) {
        std::printf("status %d: %s\n", (int)e.status, e.what());
        return 3;
    }
    return 0;
}
