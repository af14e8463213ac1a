#include "gpu_layout.hpp"

#include <cstring>
#include <stdexcept>

namespace pray {

namespace {

inline uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

struct Flattener {
    KDTreeGPU &g;

    void leaf(const KDTreeElement *n, uint32_t slot) {
        g.numLeaves++;
        if (n->triangleIndices.empty()) {
            g.numEmptyLeaves++;
            g.nodes[2 * slot + 1] = (LEAF_EMPTY << 2) | LEAF_TAG;
            return;
        }
        const uint32_t first = static_cast<uint32_t>(g.leafRefs.size());
        if (first >= LEAF_EMPTY) throw std::runtime_error("leaf reference array exceeds 2^30 entries");
        const size_t cnt = n->triangleIndices.size();
        for (size_t k = 0; k < cnt; k++) {
            const uint32_t t = n->triangleIndices[k];
            g.leafRefs.push_back(t | (k + 1 == cnt ? REF_LAST : 0u));
        }
        g.nodes[2 * slot + 1] = (first << 2) | LEAF_TAG;
    }

    // Depth-first: a node's children pair is allocated when the node is
    // visited, then the left subtree is laid out before the right one.
    void visit(const KDTreeElement *n, uint32_t slot, uint32_t depth) {
        if (depth > g.maxDepth) g.maxDepth = depth;
        if (!n->left && !n->right) {
            leaf(n, slot);
            return;
        }
        const uint32_t axis = n->splitPlane.axis;
        if (axis > 2) throw std::runtime_error("inner node without split axis");
        const uint32_t lc = static_cast<uint32_t>(g.nodes.size() / 2);
        if (lc >= LEAF_EMPTY) throw std::runtime_error("node array exceeds 2^30 entries");
        g.nodes.resize(g.nodes.size() + 4, 0u);
        g.nodes[2 * slot + 1] = (lc << 2) | axis;
        KDTreeElement emptyLeaf;
        const KDTreeElement *L = n->left ? n->left.get() : &emptyLeaf;
        const KDTreeElement *R = n->right ? n->right.get() : &emptyLeaf;
        // Clip planes along this node's axis: the left child's upper bound and
        // the right child's lower bound.  A missing child is empty, so any
        // value works; the parent's bound is used.
        const float clipL = n->left ? L->boundingBox.max[axis] : n->boundingBox.min[axis];
        const float clipR = n->right ? R->boundingBox.min[axis] : n->boundingBox.max[axis];
        g.nodes[2 * lc] = fbits(clipL);
        g.nodes[2 * (lc + 1)] = fbits(clipR);
        visit(L, lc, depth + 1);
        visit(R, lc + 1, depth + 1);
    }
};

}  // namespace

KDTreeGPU buildGPUTreeFromKDTree(const KDTree &tree) {
    KDTreeGPU g;
    const size_t n = tree.triangles.size();
    g.triA.resize(3 * n);
    g.triAB.resize(3 * n);
    g.triAC.resize(3 * n);
    g.triNormal.resize(3 * n);
    g.triColor.resize(3 * n);
    g.triEmit.resize(n);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
        const Triangle &t = tree.triangles[i];
        const float a[3] = {t.a.x, t.a.y, t.a.z}, ab[3] = {t.ab.x, t.ab.y, t.ab.z}, ac[3] = {t.ac.x, t.ac.y, t.ac.z};
        const float nn[3] = {t.normal.x, t.normal.y, t.normal.z}, c[3] = {t.color.x, t.color.y, t.color.z};
        for (int k = 0; k < 3; k++) {
            g.triA[3 * i + k] = a[k];
            g.triAB[3 * i + k] = ab[k];
            g.triAC[3 * i + k] = ac[k];
            g.triNormal[3 * i + k] = nn[k];
            g.triColor[3 * i + k] = c[k];
        }
        g.triEmit[i] = t.isEmitting ? 1 : 0;
    }
    if (n == 0 || !tree.root) return g;  // empty scene: no nodes

    g.nodes.assign(4, 0u);  // node 0 (padding) and node 1 (root)
    g.leafRefs.reserve(n + n / 2);
    Flattener f{g};
    f.visit(tree.root.get(), 1, 0);
    const BoundingBox &rb = tree.root->boundingBox;
    g.rootBox[0] = rb.min.x, g.rootBox[1] = rb.min.y, g.rootBox[2] = rb.min.z;
    g.rootBox[3] = rb.max.x, g.rootBox[4] = rb.max.y, g.rootBox[5] = rb.max.z;
    return g;
}

}  // namespace pray
