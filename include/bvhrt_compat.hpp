// bvhrt_compat.hpp -- header-only C++ facade with the BVHRT-Core builder API
// over the C ABI in atrbvh.h, so the reference's call sequence
// (gpu-ray-traversal/src/rt/cuda/Renderer.cpp:164-216) compiles with only
// its types swapped.  Reference classes mirrored:
//   SceneWrapper            BVHRT-Core/src/SceneWrapper.h:25-45
//   LBVHBuilder             BVHRT-Core/src/LBVHBuilder.h:24-65
//   TRBVHOptimizer          BVHRT-Core/src/TRBVHOptimizer.h:20-46
//   AgglomerativeTreeletOptimizer BVHRT-Core/src/AgglomerativeTreeletOptimizer.h:17-42
//   BVHTreeCollapser        BVHRT-Core/src/BVHTreeCollapser.h:14-41
//   BVHTreeInstanceManager  BVHRT-Core/src/BVHTreeInstanceManager.h:18-133
//   SoABVHTree / BVHCollapsedTree host mirrors (SoABVHTree.h, BVHCollapsedTree.h)
// Errors: the reference prints and exits (CudaErrorCheck.h:12-23); this facade
// throws BVHRT::Error carrying the atr_status and message instead.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "atrbvh.h"

namespace BVHRT {

struct Error : std::runtime_error {
    atr_status status;
    Error(atr_status s, const std::string& what) : std::runtime_error(what), status(s) {}
};

inline void check(atr_status s)
{
    if (s != ATR_OK) throw Error(s, atr_last_error());
}

struct Float3 {
    float x, y, z;
};
struct Float4 {
    float x, y, z, w;
};

// One context per process / device, created on first use (the reference runs
// on the implicit CUDA context and legacy stream).
inline atr_context* DefaultContext(int device = 0)
{
    static atr_context* ctx = nullptr;
    if (!ctx) check(atr_context_create(device, &ctx));
    return ctx;
}

// SceneWrapper(numberOfTriangles, float4* vertices, bbMin, bbMax): copies the
// vertices to the device.  Unlike the reference it does not take ownership of
// the host array.
class SceneWrapper {
public:
    SceneWrapper(unsigned int numberOfTriangles, const Float4* vertices, Float3 bbMin,
                 Float3 bbMax, atr_context* ctx = DefaultContext())
    {
        const float mn[3] = {bbMin.x, bbMin.y, bbMin.z};
        const float mx[3] = {bbMax.x, bbMax.y, bbMax.z};
        check(atr_scene_create(ctx, reinterpret_cast<const float*>(vertices), numberOfTriangles,
                               mn, mx, &scene_));
    }
    // Scene box computed on the device with SceneConverter's rule.
    SceneWrapper(unsigned int numberOfTriangles, const Float4* vertices,
                 atr_context* ctx = DefaultContext())
    {
        check(atr_scene_create(ctx, reinterpret_cast<const float*>(vertices), numberOfTriangles,
                               nullptr, nullptr, &scene_));
    }
    ~SceneWrapper() { atr_scene_destroy(scene_); }
    SceneWrapper(const SceneWrapper&) = delete;
    SceneWrapper& operator=(const SceneWrapper&) = delete;
    atr_scene* DeviceScene() const { return scene_; }

private:
    atr_scene* scene_ = nullptr;
};

// A device tree (BVHTree* in the reference, owned by the caller and freed by
// BVHTreeInstanceManager::FreeDeviceTree).
using BVHTree = atr_tree;
using DeviceCollapsedTree = atr_collapsed;

class BVHBuilder {
public:
    virtual ~BVHBuilder() = default;
    virtual BVHTree* BuildTree(const SceneWrapper* scene) = 0;
};

class LBVHBuilder : public BVHBuilder {
public:
    explicit LBVHBuilder(bool use64BitMortonCodes = false) : bits_(use64BitMortonCodes ? 63 : 30) {}
    BVHTree* BuildTree(const SceneWrapper* scene) override
    {
        atr_tree* t = nullptr;
        check(atr_lbvh_build(scene->DeviceScene(), bits_, &t));
        return t;
    }

private:
    int bits_;
};

class BVHOptimizer {
public:
    virtual ~BVHOptimizer() = default;
    virtual void Optimize(BVHTree* deviceTree) = 0;
};

class TRBVHOptimizer : public BVHOptimizer {
public:
    TRBVHOptimizer(int treeletSize = 7, int iterations = 3) : n_(treeletSize), it_(iterations) {}
    void Optimize(BVHTree* t) override { check(atr_trbvh_optimize(t, n_, it_, 1.2f, 1.0f)); }

private:
    int n_, it_;
};

class AgglomerativeTreeletOptimizer : public BVHOptimizer {
public:
    AgglomerativeTreeletOptimizer(int treeletSize = 9, int iterations = 2)
        : n_(treeletSize), it_(iterations)
    {
    }
    void Optimize(BVHTree* t) override { check(atr_atrbvh_optimize(t, n_, it_, 1.2f, 1.0f)); }

private:
    int n_, it_;
};

// Host mirrors (DeviceToHostTree / DeviceToHostCollapsedTree results).
struct SoABVHTree {
    unsigned int numberOfTriangles = 0;
    int rootIndex = 0;
    std::vector<int> parentIndices, leftIndices, rightIndices, dataIndices;
    std::vector<Float4> boundingBoxMin, boundingBoxMax;
    std::vector<float> area;
    unsigned int NumberOfTriangles() const { return numberOfTriangles; }
    int RootIndex() const { return rootIndex; }
    int ParentIndex(int i) const { return parentIndices[i]; }
    int LeftIndex(int i) const { return leftIndices[i]; }
    int RightIndex(int i) const { return rightIndices[i]; }
    int DataIndex(int i) const { return dataIndices[i]; }
    Float4 BoundingBoxMin(int i) const { return boundingBoxMin[i]; }
    Float4 BoundingBoxMax(int i) const { return boundingBoxMax[i]; }
    float Area(int i) const { return area[i]; }
    // SoABVHTree::SAH (SoABVHTree.cpp:50-74): same summation order
    float SAH(float ci = 1.2f, float ct = 1.0f) const
    {
        const float total = area[rootIndex];
        float si = total, sl = 0.0f;
        for (unsigned int i = 0; i + 1 < numberOfTriangles; ++i)
            if ((int)i != rootIndex) si += area[i];
        for (unsigned int i = numberOfTriangles - 1; i < 2 * numberOfTriangles - 1; ++i)
            sl += area[i];
        return (ci * si + ct * sl) / total;
    }
};

struct BVHCollapsedTree {
    unsigned int numberOfTriangles = 0;
    int rootIndex = 0;
    unsigned int used = 0;
    std::vector<int> triangleIndices;
    struct {
        std::vector<int> parentIndices, leftIndices, rightIndices, dataIndices, triangleCount;
        std::vector<Float4> boundingBoxMin, boundingBoxMax;
    } nodes;
};

class BVHTreeCollapser {
public:
    DeviceCollapsedTree* Collapse(BVHTree* deviceTree, float* sah, float ci = 1.2f,
                                  float ct = 1.0f)
    {
        atr_collapsed* c = nullptr;
        check(atr_collapse(deviceTree, ci, ct, &c, sah));
        return c;
    }
};

class BVHTreeInstanceManager {
public:
    SoABVHTree* DeviceToHostTree(BVHTree* t) const
    {
        uint32_t n = 0;
        check(atr_tree_num_triangles(t, &n));
        SoABVHTree* h = new SoABVHTree();
        const size_t m = 2ull * n - 1;
        h->numberOfTriangles = n;
        h->parentIndices.resize(m);
        h->leftIndices.resize(m);
        h->rightIndices.resize(m);
        h->dataIndices.resize(m);
        h->boundingBoxMin.resize(m);
        h->boundingBoxMax.resize(m);
        h->area.resize(m);
        atr_tree_view v{n,
                        0,
                        h->parentIndices.data(),
                        h->leftIndices.data(),
                        h->rightIndices.data(),
                        h->dataIndices.data(),
                        &h->boundingBoxMin[0].x,
                        &h->boundingBoxMax[0].x,
                        h->area.data()};
        check(atr_tree_download(t, &v));
        h->rootIndex = v.root;
        return h;
    }
    BVHCollapsedTree* DeviceToHostCollapsedTree(DeviceCollapsedTree* c) const
    {
        uint32_t n = 0;
        check(atr_collapsed_num_triangles(c, &n));
        BVHCollapsedTree* h = new BVHCollapsedTree();
        const size_t m = 2ull * n - 1;
        h->numberOfTriangles = n;
        h->nodes.parentIndices.resize(m);
        h->nodes.leftIndices.resize(m);
        h->nodes.rightIndices.resize(m);
        h->nodes.dataIndices.resize(m);
        h->nodes.triangleCount.resize(m);
        h->nodes.boundingBoxMin.resize(m);
        h->nodes.boundingBoxMax.resize(m);
        h->triangleIndices.resize(n);
        atr_collapsed_view v{n,
                             0,
                             0,
                             h->nodes.parentIndices.data(),
                             h->nodes.leftIndices.data(),
                             h->nodes.rightIndices.data(),
                             h->nodes.dataIndices.data(),
                             h->nodes.triangleCount.data(),
                             &h->nodes.boundingBoxMin[0].x,
                             &h->nodes.boundingBoxMax[0].x,
                             h->triangleIndices.data()};
        check(atr_collapsed_download(c, &v));
        h->rootIndex = v.root;
        h->used = v.used;
        return h;
    }
    void FreeDeviceTree(BVHTree* t) const { atr_tree_destroy(t); }
    void FreeDeviceCollapsedTree(DeviceCollapsedTree* c) const { atr_collapsed_destroy(c); }
};

}  // namespace BVHRT
