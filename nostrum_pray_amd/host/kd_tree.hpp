// CPU KD-tree builders of the drop-in host (they stay on the CPU).
//
// KDTree     -- object-median split on the longest axis of the node's
//               triangle-box union (reference src/kd_tree.cpp:25-84,
//               include/kd_tree.h:49-52).  Child boxes may overlap.
// KDTreeSAH  -- Wald-Havran event-sweep SAH over disjoint spatial cells
//               (reference src/kd_tree_sah.cpp:32-296,
//               include/kd_tree_sah.h:14-91).
// KDTreeBinnedSAH -- the SAH cost on 64 bins per axis (this build's
//               extension, the CPU twin of the GPU build).
//
// Both produce the same pointer tree type (KDTreeElement, mirroring
// include/kd_tree.h:80-134).  Leaf triangle lists, split planes and node
// boxes are identical to the reference builders' for the same input (the
// tests compare them leaf by leaf against the reference compiled from
// source).  Triangle indices are 32-bit here.  There is no CPU ray traversal
// in this library: rays are traced by the GPU through include/pray_hip.h.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <vector>

#include "geometry.hpp"

namespace pray {

typedef enum { LEFT = -1, RIGHT = 1, UNKNOWN = 0 } PlaneSide;
typedef enum { LEFT_ONLY = -1, RIGHT_ONLY = 1, BOTH = 0 } TriangleSide;

struct KDTreeElement {
    BoundingBox boundingBox;
    std::vector<uint32_t> triangleIndices;
    std::unique_ptr<KDTreeElement> left, right;
    SplitPlane splitPlane;
    PlaneSide pSide = UNKNOWN;

    // Subtree statistics gathered bottom-up (kd_tree.h:124-133).
    size_t nSubNodes = 0, maxDepthToLeaf = 0, nSubLeafs = 0, nSubEmptyLeafs = 0;

    KDTreeElement() : boundingBox(Vec3(0.f, 0.f, 0.f), Vec3(0.f, 0.f, 0.f)) {}
    KDTreeElement(const BoundingBox &b, std::vector<uint32_t> &&tris) : boundingBox(b), triangleIndices(std::move(tris)) {}

    bool isLeaf() const { return !triangleIndices.empty() || (!left && !right); }
    void log();
};

// Interface of the spatial database (reference include/spatial.h:10-39),
// without the CPU intersect methods: intersection is the GPU's job.
class SpatialDatabase {
public:
    virtual ~SpatialDatabase() = default;
    virtual void build(const std::vector<Triangle> &triangles) = 0;
    virtual void build(std::vector<Triangle> &&triangles) = 0;
    virtual bool isEmpty() const = 0;
    virtual size_t size() const = 0;

    size_t maxdepth = 0, nnodes = 0, leafNodes = 0, emptyLeafNodes = 0, triangleCount = 0;
};

class KDTree : public SpatialDatabase {
public:
    void build(const std::vector<Triangle> &triangles) override;
    void build(std::vector<Triangle> &&triangles) override;
    bool isEmpty() const override { return triangles.empty(); }
    size_t size() const override { return triangles.size(); }

    std::unique_ptr<KDTreeElement> root;
    std::vector<Triangle> triangles;

    static constexpr size_t THRESHOLD = 32;   // kd_tree.h:50
    static constexpr size_t MAX_DEPTH = 20;   // kd_tree.h:52
    // Depth cap used by the build; MAX_DEPTH reproduces the reference tree.
    // A larger cap is an extension of this build (deeper trees, smaller leaves).
    size_t depthLimit = MAX_DEPTH;
    // Leaf size below which a node is never split; THRESHOLD reproduces the
    // reference tree, a smaller one is an extension (smaller leaves).
    size_t leafSize = THRESHOLD;
    // Parallel task depth; results do not depend on it.
    static constexpr size_t TASK_DEPTH = 8;

protected:
    void finishStats();

private:
    void buildMedian();
    std::unique_ptr<KDTreeElement> buildRecursive(std::vector<uint32_t> &triIndices, size_t depth);
};

class KDTreeSAH : public KDTree {
public:
    void build(const std::vector<Triangle> &triangles) override;
    void build(std::vector<Triangle> &&triangles) override;

    // SAH constants (kd_tree_sah.h:49-50).  Reference values by default;
    // other values are an extension for traversal hardware with other costs.
    static constexpr float REF_TRAVERSAL_COST = 5.f;
    static constexpr float REF_INTERSECTION_COST = 1.5f;
    float traversalCost = REF_TRAVERSAL_COST;
    float intersectionCost = REF_INTERSECTION_COST;

private:
    void buildSAH();
};

// Binned-SAH KD tree (no reference counterpart; an extension of this build):
// the reference's SAH cost (traversal + intersection x area-weighted counts,
// 0.8 for an empty side, kd_tree_sah.cpp:141-166) evaluated on the 63
// boundaries of 64 equal bins per axis of the node's cell instead of on every
// triangle event.  A reference goes left when its box clipped to the cell
// starts below the plane (or lies in it) and right when it ends above it.
// The decisions and their floating-point steps are those of the GPU build
// (pray_build_scene with PRAY_BUILD_SAH), so both give the same tree; this
// one is flattened with the host's treelet layout.
class KDTreeBinnedSAH : public KDTree {
public:
    void build(const std::vector<Triangle> &triangles) override;
    void build(std::vector<Triangle> &&triangles) override;

    static constexpr int BINS = 64;
    float traversalCost = 2.f;
    float intersectionCost = 1.f;

private:
    void buildBinned();
    std::unique_ptr<KDTreeElement> rec(std::vector<uint32_t> &&refs, const BoundingBox &cell, size_t depth);
};

// Export of a built tree in depth-first, left-before-right order, used to
// compare trees with other builders: for each node one record
// {is_leaf, axis, split, box[6], first, count} and the concatenated leaf
// triangle lists.
struct TreeExport {
    std::vector<int32_t> is_leaf, axis;
    std::vector<float> split, box;
    std::vector<uint32_t> first, count, tris;
};
void exportTree(const KDTreeElement *root, TreeExport &out);

}  // namespace pray
