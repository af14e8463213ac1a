// Flat device layout of a KD tree: 8-byte nodes, leaf triangle-index arrays
// and per-field triangle arrays.  Replaces the reference's 88-byte
// KDTreeGPUNode / KDTreeGPU (include/kd_tree_gpu_structs.h:7-22,
// include/kd_tree_gpu.h:11-19) and its BFS flattening
// (src/gpu_util.cpp:7-57).
//
// Node (8 bytes, two uint32 words):
//   word0  float bits of the node's CLIP plane along its PARENT's split axis:
//          for a left child, the child's upper bound; for a right child, its
//          lower bound.  (Unused for the root.)
//   word1  bits[1:0] = split axis 0/1/2 of this node, or 3 for a leaf.
//          inner: bits[31:2] = index of the left child; the right child is the
//                 next node.  Children pairs start at even indices, so a pair
//                 is one aligned 16-byte load.
//          leaf:  bits[31:2] = index of the first entry of the leaf's run in
//                 leaf_refs, or LEAF_EMPTY.
// The root is node 1 (node 0 is padding).  Nodes are stored in 128-byte
// blocks of 8 children pairs: a node's children pair and, breadth first, the
// pairs of its descendants share a block until it is full (a treelet), so one
// cache line serves up to three levels of descent.  Unused slots are zero.
//
// leaf_refs: uint32 per (leaf, triangle) reference: triangle index in bits
//   [30:0]; bit 31 marks the last entry of a leaf's run.
//
// For the SAH tree both children's clips equal the split (disjoint cells);
// for the median tree they are the children's box bounds along the split axis
// (bounding-interval form; children may overlap).  The same traversal serves
// both.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "kd_tree.hpp"

namespace pray {

constexpr uint32_t LEAF_TAG = 3u;
constexpr uint32_t LEAF_EMPTY = 0x3FFFFFFFu;
constexpr uint32_t REF_LAST = 0x80000000u;

struct KDTreeGPU {
    std::vector<uint32_t> nodes;      // 2 words per node
    std::vector<uint32_t> leafRefs;   // see above
    std::vector<float> triA, triAB, triAC;               // 3 floats per triangle each
    std::vector<float> triNormal, triColor;              // 3 floats per triangle each
    std::vector<uint8_t> triEmit;                         // 1 per triangle
    float rootBox[6] = {0, 0, 0, 0, 0, 0};                // min xyz, max xyz
    uint32_t maxDepth = 0;                                // root-to-leaf edges
    uint32_t numLeaves = 0, numEmptyLeaves = 0;

    uint32_t numNodes() const { return static_cast<uint32_t>(nodes.size() / 2); }
    uint32_t numTris() const { return static_cast<uint32_t>(triEmit.size()); }
};

// Flattens a built tree.  Unlike the reference (gpu_util.cpp:13), the CPU tree
// keeps its triangles; they are copied into the per-field arrays.
KDTreeGPU buildGPUTreeFromKDTree(const KDTree &tree);

// A flattened tree as a file (flat_io.cpp): written once, uploaded later
// without rebuilding.  The load checks the header, the length and a
// checksum; the upload validates the structure.
bool saveFlat(const KDTreeGPU &g, int kind, const std::string &path, std::string &err);
bool loadFlat(const std::string &path, KDTreeGPU &g, int &kind, std::string &err);

}  // namespace pray
