#include "gpu_layout.hpp"

#include <cstring>
#include <stdexcept>
#include <unordered_map>

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

    // Writes inner node n (at `slot`) with its children pair at node index lc.
    void placePair(const KDTreeElement *n, uint32_t slot, uint32_t lc) {
        const uint32_t axis = n->splitPlane.axis;
        if (axis > 2) throw std::runtime_error("inner node without split axis");
        g.nodes[2 * slot + 1] = (lc << 2) | axis;
        // Clip planes along this node's axis: the left child's upper bound and
        // the right child's lower bound.  A missing child is empty, so any
        // value works; the parent's bound is used.
        const float clipL = n->left ? n->left->boundingBox.max[axis] : n->boundingBox.min[axis];
        const float clipR = n->right ? n->right->boundingBox.min[axis] : n->boundingBox.max[axis];
        g.nodes[2 * lc] = fbits(clipL);
        g.nodes[2 * (lc + 1)] = fbits(clipR);
    }

    // A fresh block of 16 nodes (8 children pairs, 128 bytes); returns its
    // first node index.
    uint32_t newBlock() {
        const size_t base = g.nodes.size() / 2;
        if (base + 16 >= LEAF_EMPTY) throw std::runtime_error("node array exceeds 2^30 entries");
        g.nodes.resize(g.nodes.size() + 32, 0u);
        return static_cast<uint32_t>(base);
    }

    struct Item {
        const KDTreeElement *n;
        uint32_t slot, depth;
    };

    std::unordered_map<const KDTreeElement *, uint32_t> innerCount;
    uint32_t packBase = 0, packNext = 8;  // shared block for small subtrees

    static bool isLeaf(const KDTreeElement *n) { return !n->left && !n->right; }

    uint32_t countInner(const KDTreeElement *n) {
        if (!n || isLeaf(n)) return 0;
        const uint32_t c = 1 + countInner(n->left.get()) + countInner(n->right.get());
        innerCount[n] = c;
        return c;
    }

    // Places the children pairs of `root` and of its inner descendants
    // breadth first into [base + 2*next, base + 16) until `limit` pairs are
    // placed; returns the inner nodes left over (their pairs not placed).
    std::vector<Item> fill(Item root, uint32_t base, uint32_t &next, uint32_t limit) {
        KDTreeElement emptyLeaf;
        std::vector<Item> queue{root}, rest;
        for (size_t qi = 0; qi < queue.size(); qi++) {
            const Item it = queue[qi];
            if (next >= limit) {
                rest.push_back(it);
                continue;
            }
            const uint32_t lc = base + 2 * next++;
            placePair(it.n, it.slot, lc);
            const KDTreeElement *kids[2] = {it.n->left ? it.n->left.get() : &emptyLeaf,
                                            it.n->right ? it.n->right.get() : &emptyLeaf};
            for (uint32_t k = 0; k < 2; k++) {
                const uint32_t d = it.depth + 1;
                if (d > g.maxDepth) g.maxDepth = d;
                if (isLeaf(kids[k]))
                    leaf(kids[k], lc + k);
                else
                    queue.push_back({kids[k], lc + k, d});
            }
        }
        return rest;
    }

    // Treelet layout: an inner node's children pair and, breadth first, the
    // pairs of its descendants fill one 128-byte block of 8 pairs; inner
    // nodes left over start blocks of their own.  A subtree of at most 8
    // inner nodes is placed whole into a block shared with other small
    // subtrees (first fit).  One cache line thus serves up to three levels of
    // descent, and children pairs always lie after their parent.
    void place(Item it) {
        const uint32_t c = innerCount[it.n];
        if (c <= 8) {
            if (8 - packNext < c) {
                packBase = newBlock();
                packNext = 0;
            }
            const uint32_t limit = packNext + c;
            fill(it, packBase, packNext, limit);
            return;
        }
        uint32_t next = 0;
        for (const Item &r : fill(it, newBlock(), next, 8)) place(r);
    }

    void layout(const KDTreeElement *root) {
        // Block 0 holds the padding node 0 and the root (node 1) in its
        // first pair slot.
        const uint32_t base = newBlock();
        if (isLeaf(root)) {
            leaf(root, 1);
            return;
        }
        countInner(root);
        uint32_t next = 1;
        for (const Item &r : fill({root, 1, 0}, base, next, 8)) place(r);
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

    g.nodes.clear();
    g.leafRefs.reserve(n + n / 2);
    Flattener f{g};
    f.layout(tree.root.get());
    const BoundingBox &rb = tree.root->boundingBox;
    g.rootBox[0] = rb.min.x, g.rootBox[1] = rb.min.y, g.rootBox[2] = rb.min.z;
    g.rootBox[3] = rb.max.x, g.rootBox[4] = rb.max.y, g.rootBox[5] = rb.max.z;
    return g;
}

}  // namespace pray
