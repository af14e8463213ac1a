// CPU KD-tree builders (median and SAH).  See kd_tree.hpp for the contract.
#include "kd_tree.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <parallel/algorithm>

namespace pray {

void KDTreeElement::log() {
    nSubNodes = left->nSubNodes + right->nSubNodes + 2;
    maxDepthToLeaf = std::max(left->maxDepthToLeaf, right->maxDepthToLeaf) + 1;
    nSubLeafs += left->isLeaf() ? 1 : left->nSubLeafs;
    nSubLeafs += right->isLeaf() ? 1 : right->nSubLeafs;
    nSubEmptyLeafs += left->isLeaf() ? (left->triangleIndices.empty() ? 1 : 0) : left->nSubEmptyLeafs;
    nSubEmptyLeafs += right->isLeaf() ? (right->triangleIndices.empty() ? 1 : 0) : right->nSubEmptyLeafs;
}

void KDTree::finishStats() {
    nnodes = root->nSubNodes + 1;
    leafNodes = root->nSubLeafs;
    emptyLeafNodes = root->nSubEmptyLeafs;
    maxdepth = root->maxDepthToLeaf;
}

// ---------------------------------------------------------------- median ---

void KDTree::build(const std::vector<Triangle> &tris) {
    triangles = tris;
    triangleCount = triangles.size();
    buildMedian();
}

void KDTree::build(std::vector<Triangle> &&tris) {
    triangles = std::move(tris);
    triangleCount = triangles.size();
    buildMedian();
}

void KDTree::buildMedian() {
    if (triangles.empty()) {
        root = std::make_unique<KDTreeElement>();
        return;
    }
    std::vector<uint32_t> idx(triangles.size());
    std::iota(idx.begin(), idx.end(), 0u);
#pragma omp parallel
#pragma omp single nowait
    root = buildRecursive(idx, 0);
    finishStats();
}

// Object median on the longest axis of the node's box union
// (kd_tree.cpp:48-84).  std::nth_element is called with the same comparator
// on the same index order, so the partition is the reference's.
std::unique_ptr<KDTreeElement> KDTree::buildRecursive(std::vector<uint32_t> &triIndices, size_t depth) {
    BoundingBox nodeBB = triangles[triIndices[0]].boundingBox;
    for (size_t i = 1; i < triIndices.size(); i++) nodeBB = nodeBB.expand(triangles[triIndices[i]].boundingBox);
    const uint8_t axis = nodeBB.getLongestAxis();

    if (triIndices.size() <= leafSize || depth >= depthLimit) {
        std::vector<uint32_t> copy(triIndices);
        return std::make_unique<KDTreeElement>(nodeBB, std::move(copy));
    }

    const size_t half = triIndices.size() / 2;
    std::nth_element(triIndices.begin(), triIndices.begin() + half, triIndices.end(),
                     [&](uint32_t l, uint32_t r) { return triangles[l].midPoint[axis] < triangles[r].midPoint[axis]; });

    std::vector<uint32_t> leftIdx(triIndices.begin(), triIndices.begin() + half);
    std::vector<uint32_t> rightIdx(triIndices.begin() + half, triIndices.end());

    // Children as OpenMP tasks near the root, plain calls below (an undeferred
    // task would run in place anyway, at the cost of a task descriptor).
    std::unique_ptr<KDTreeElement> l, r;
    if (depth < TASK_DEPTH) {
#pragma omp task shared(l, leftIdx)
        l = buildRecursive(leftIdx, depth + 1);
#pragma omp task shared(r, rightIdx)
        r = buildRecursive(rightIdx, depth + 1);
#pragma omp taskwait
    } else {
        l = buildRecursive(leftIdx, depth + 1);
        r = buildRecursive(rightIdx, depth + 1);
    }

    auto node = std::make_unique<KDTreeElement>();
    node->boundingBox = nodeBB;
    node->left = std::move(l);
    node->right = std::move(r);
    node->splitPlane.axis = axis;
    node->splitPlane.distance = triangles[triIndices[half]].midPoint[axis];
    node->log();
    return node;
}

// ------------------------------------------------------------------- SAH ---

namespace {

enum : uint8_t { EV_END = 0, EV_LYING = 1, EV_START = 2 };

// A split candidate produced by one triangle (kd_tree_sah.h:14-36).  `slot`
// is the triangle's position in the owning node's triangle list, which
// replaces the reference's global-index std::map with a flat array.
struct SahEvent {
    float dist;
    uint32_t slot;
    uint8_t axis;
    uint8_t type;
};

// Order (distance, axis, type): the reference's Event::operator<.
inline bool evLess(const SahEvent &a, const SahEvent &b) {
    return (a.dist < b.dist) || (a.dist == b.dist && a.axis < b.axis) ||
           (a.dist == b.dist && a.axis == b.axis && a.type < b.type);
}

// Events of one triangle's box clipped to a voxel (kd_tree_sah.cpp:272-284).
inline void genEvents(const Triangle &tri, const BoundingBox &box, uint32_t slot, std::vector<SahEvent> &out) {
    const BoundingBox eb = tri.boundingBox.clip(box);
    for (uint8_t k = 0; k < 3; k++) {
        if (eb.isPlanar(k)) {
            out.push_back({eb.min[k], slot, k, EV_LYING});
        } else {
            out.push_back({eb.min[k], slot, k, EV_START});
            out.push_back({eb.max[k], slot, k, EV_END});
        }
    }
}

inline float lambdaBias(size_t nl, size_t nr, float pl, float pr) {
    if ((nl == 0 || nr == 0) && !(pl == 1 || pr == 1)) return 0.8f;
    return 1.0f;
}

struct SahCost {
    float traversal, intersection;
};

inline float costOf(const SahCost &c, float pl, float pr, size_t nl, size_t nr) {
    return lambdaBias(nl, nr, pl, pr) * (c.traversal + c.intersection * (pl * nl + pr * nr));
}

// SAH cost of one plane (kd_tree_sah.cpp:141-166).
void sah(const SahCost &c, const SplitPlane &p, const BoundingBox &b, size_t nl, size_t nr, size_t np, float &costP,
         PlaneSide &pside) {
    costP = INFTY;
    BoundingBox bl, br;
    b.split(p, bl, br);
    const float pl = bl.surfaceArea() / b.surfaceArea();
    const float pr = br.surfaceArea() / b.surfaceArea();
    if (pl == 0 || pr == 0) return;
    if (b.max[p.axis] - b.min[p.axis] == 0) return;
    const float costPl = costOf(c, pl, pr, nl + np, nr);
    const float costPr = costOf(c, pl, pr, nl, np + nr);
    if (costPl < costPr) {
        costP = costPl;
        pside = LEFT;
    } else {
        costP = costPr;
        pside = RIGHT;
    }
}

// Sweep over the sorted events of a node (kd_tree_sah.cpp:169-207).  The
// events of one plane are consumed by three while-loops; the enclosing for's
// increment then steps once more, so the first event after each plane group is
// never counted.  That behaviour decides which planes and counts the reference
// evaluates, so it is kept.
void findPlane(const SahCost &cost, size_t numTriangles, const std::vector<SahEvent> &ev, const BoundingBox &b,
               SplitPlane &pest, float &cest, PlaneSide &psideest) {
    cest = INFTY;
    size_t Nl[3] = {0, 0, 0}, Np[3] = {0, 0, 0}, Nr[3] = {numTriangles, numTriangles, numTriangles};
    const size_t n = ev.size();
    for (size_t i = 0; i < n; i++) {
        const float pd = ev[i].dist;
        const uint8_t pa = ev[i].axis;
        size_t pLying = 0, pStarting = 0, pEnding = 0;
        while (i < n && ev[i].axis == pa && ev[i].dist == pd && ev[i].type == EV_END) {
            pEnding++;
            i++;
        }
        while (i < n && ev[i].axis == pa && ev[i].dist == pd && ev[i].type == EV_LYING) {
            pLying++;
            i++;
        }
        while (i < n && ev[i].axis == pa && ev[i].dist == pd && ev[i].type == EV_START) {
            pStarting++;
            i++;
        }
        Np[pa] = pLying;
        Nr[pa] -= pLying;
        Nr[pa] -= pEnding;

        float C;
        PlaneSide ps = UNKNOWN;
        const SplitPlane p(pa, pd);
        sah(cost, p, b, Nl[pa], Nr[pa], Np[pa], C, ps);
        if (C < cest) {
            cest = C;
            pest = p;
            psideest = ps;
        }
        Nl[pa] += pStarting;
        Nl[pa] += pLying;
        Np[pa] = 0;
    }
}

struct SahBuilder {
    const std::vector<Triangle> &tris;
    size_t depthLimit, leafSize;
    SahCost cost;

    std::unique_ptr<KDTreeElement> rec(std::vector<uint32_t> &T, const BoundingBox &b, std::vector<SahEvent> &E,
                                       const SplitPlane &prevp, size_t depth) {
        SplitPlane p;
        float Cp;
        PlaneSide pside = UNKNOWN;
        findPlane(cost, T.size(), E, b, p, Cp, pside);

        if (T.size() <= leafSize || depth >= depthLimit || p == prevp || Cp > cost.intersection * T.size()) {
            std::vector<SahEvent>().swap(E);
            return std::make_unique<KDTreeElement>(b, std::move(T));
        }

        // Classification (kd_tree_sah.cpp:209-233).
        std::vector<int8_t> side(T.size(), BOTH);
        for (const SahEvent &e : E) {
            if (e.axis != p.axis) continue;
            if (e.type == EV_END && e.dist <= p.distance) {
                side[e.slot] = LEFT_ONLY;
            } else if (e.type == EV_START && e.dist >= p.distance) {
                side[e.slot] = RIGHT_ONLY;
            } else if (e.type == EV_LYING) {
                if (e.dist < p.distance || (e.dist == p.distance && pside == LEFT))
                    side[e.slot] = LEFT_ONLY;
                else if (e.dist > p.distance || (e.dist == p.distance && pside == RIGHT))
                    side[e.slot] = RIGHT_ONLY;
            }
        }

        BoundingBox VL, VR;
        b.split(p, VL, VR);

        // Child triangle lists in parent order; straddlers go to both sides
        // with events regenerated against each child voxel
        // (kd_tree_sah.cpp:250-270).
        std::vector<uint32_t> tl, tr, lslot(T.size()), rslot(T.size());
        std::vector<SahEvent> bl, br;
        for (size_t s = 0; s < T.size(); s++) {
            if (side[s] == LEFT_ONLY) {
                lslot[s] = static_cast<uint32_t>(tl.size());
                tl.push_back(T[s]);
            } else if (side[s] == RIGHT_ONLY) {
                rslot[s] = static_cast<uint32_t>(tr.size());
                tr.push_back(T[s]);
            } else {
                genEvents(tris[T[s]], VL, static_cast<uint32_t>(tl.size()), bl);
                genEvents(tris[T[s]], VR, static_cast<uint32_t>(tr.size()), br);
                tl.push_back(T[s]);
                tr.push_back(T[s]);
            }
        }

        // One-sided events keep their sorted order (kd_tree_sah.cpp:235-248).
        std::vector<SahEvent> lo, ro;
        for (const SahEvent &e : E) {
            if (side[e.slot] == LEFT_ONLY)
                lo.push_back({e.dist, lslot[e.slot], e.axis, e.type});
            else if (side[e.slot] == RIGHT_ONLY)
                ro.push_back({e.dist, rslot[e.slot], e.axis, e.type});
        }
        std::vector<SahEvent>().swap(E);
        std::vector<int8_t>().swap(side);
        std::vector<uint32_t>().swap(lslot);
        std::vector<uint32_t>().swap(rslot);

        std::sort(bl.begin(), bl.end(), evLess);
        std::sort(br.begin(), br.end(), evLess);
        std::vector<SahEvent> le(lo.size() + bl.size()), re(ro.size() + br.size());
        std::merge(lo.begin(), lo.end(), bl.begin(), bl.end(), le.begin(), evLess);
        std::merge(ro.begin(), ro.end(), br.begin(), br.end(), re.begin(), evLess);
        std::vector<SahEvent>().swap(lo);
        std::vector<SahEvent>().swap(ro);
        std::vector<SahEvent>().swap(bl);
        std::vector<SahEvent>().swap(br);

        std::unique_ptr<KDTreeElement> l, r;
        if (depth < KDTree::TASK_DEPTH) {
#pragma omp task shared(l, tl, le, VL, p)
            l = rec(tl, VL, le, p, depth + 1);
#pragma omp task shared(r, tr, re, VR, p)
            r = rec(tr, VR, re, p, depth + 1);
#pragma omp taskwait
        } else {
            l = rec(tl, VL, le, p, depth + 1);
            r = rec(tr, VR, re, p, depth + 1);
        }

        auto node = std::make_unique<KDTreeElement>();
        node->boundingBox = b;
        node->left = std::move(l);
        node->right = std::move(r);
        node->splitPlane = p;
        node->pSide = pside;
        node->log();
        return node;
    }
};

}  // namespace

void KDTreeSAH::build(const std::vector<Triangle> &tris) {
    triangles = tris;
    triangleCount = triangles.size();
    buildSAH();
}

void KDTreeSAH::build(std::vector<Triangle> &&tris) {
    triangles = std::move(tris);
    triangleCount = triangles.size();
    buildSAH();
}

void KDTreeSAH::buildSAH() {
    if (triangles.empty()) {
        root = std::make_unique<KDTreeElement>();
        return;
    }
    const size_t n = triangles.size();
    std::vector<uint32_t> T(n);
    std::iota(T.begin(), T.end(), 0u);
    BoundingBox initialBB = triangles[0].boundingBox;
    for (size_t i = 0; i < n; i++) initialBB = initialBB.expand(triangles[i].boundingBox);

    // Events of all triangles against the scene box, generated per thread in
    // index blocks and concatenated; the sort below fixes their order.
    std::vector<SahEvent> events;
    {
        const int nt = omp_get_max_threads();
        std::vector<std::vector<SahEvent>> parts(nt);
#pragma omp parallel num_threads(nt)
        {
            const int t = omp_get_thread_num();
            const size_t lo = n * t / nt, hi = n * (t + 1) / nt;
            parts[t].reserve((hi - lo) * 6);
            for (size_t i = lo; i < hi; i++) genEvents(triangles[i], initialBB, static_cast<uint32_t>(i), parts[t]);
        }
        size_t total = 0;
        for (auto &pp : parts) total += pp.size();
        events.reserve(total);
        for (auto &pp : parts) {
            events.insert(events.end(), pp.begin(), pp.end());
            std::vector<SahEvent>().swap(pp);
        }
    }
    __gnu_parallel::sort(events.begin(), events.end(), evLess);

    SahBuilder builder{triangles, depthLimit, leafSize, {traversalCost, intersectionCost}};
#pragma omp parallel
#pragma omp single nowait
    root = builder.rec(T, initialBB, events, SplitPlane(), 0);
    finishStats();
}

// ---------------------------------------------------------------- export ---

void exportTree(const KDTreeElement *root, TreeExport &out) {
    if (!root) return;
    std::vector<const KDTreeElement *> stack{root};
    while (!stack.empty()) {
        const KDTreeElement *n = stack.back();
        stack.pop_back();
        const bool leaf = !n->left && !n->right;
        out.is_leaf.push_back(leaf ? 1 : 0);
        out.axis.push_back(n->splitPlane.axis);
        out.split.push_back(n->splitPlane.distance);
        const float bx[6] = {n->boundingBox.min.x, n->boundingBox.min.y, n->boundingBox.min.z,
                             n->boundingBox.max.x, n->boundingBox.max.y, n->boundingBox.max.z};
        out.box.insert(out.box.end(), bx, bx + 6);
        out.first.push_back(static_cast<uint32_t>(out.tris.size()));
        out.count.push_back(static_cast<uint32_t>(n->triangleIndices.size()));
        out.tris.insert(out.tris.end(), n->triangleIndices.begin(), n->triangleIndices.end());
        if (!leaf) {
            if (n->right) stack.push_back(n->right.get());
            if (n->left) stack.push_back(n->left.get());
        }
    }
}


// ---------------------------------------------------------- binned SAH ---

void KDTreeBinnedSAH::build(const std::vector<Triangle> &tris) {
    triangles = tris;
    triangleCount = triangles.size();
    buildBinned();
}

void KDTreeBinnedSAH::build(std::vector<Triangle> &&tris) {
    triangles = std::move(tris);
    triangleCount = triangles.size();
    buildBinned();
}

void KDTreeBinnedSAH::buildBinned() {
    if (triangles.empty()) {
        root = std::make_unique<KDTreeElement>();
        return;
    }
    std::vector<uint32_t> T(triangles.size());
    std::iota(T.begin(), T.end(), 0u);
    BoundingBox cell = triangles[0].boundingBox;
    for (size_t i = 1; i < triangles.size(); i++) cell = cell.expand(triangles[i].boundingBox);
#pragma omp parallel
#pragma omp single nowait
    root = rec(std::move(T), cell, 0);
    finishStats();
}

namespace {
inline float halfArea(float dx, float dy, float dz) { return dx * dy + dy * dz + dz * dx; }
inline int binOf(float x, float lo, float scale) {
    return (int)std::fmax(0.f, std::fmin((float)(KDTreeBinnedSAH::BINS - 1), std::floor((x - lo) * scale)));
}
}  // namespace

std::unique_ptr<KDTreeElement> KDTreeBinnedSAH::rec(std::vector<uint32_t> &&refs, const BoundingBox &cell,
                                                    size_t depth) {
    constexpr int B = BINS;
    const float d[3] = {cell.max.x - cell.min.x, cell.max.y - cell.min.y, cell.max.z - cell.min.z};
    const float sa = halfArea(d[0], d[1], d[2]);
    const size_t n = refs.size();
    int bestA = -1;
    float bestP = 0.f;
    if (n > leafSize && depth < depthLimit && sa > 0.f) {
        uint32_t H[3][2][B] = {};
        for (uint32_t t : refs) {
            const BoundingBox &tb = triangles[t].boundingBox;
            for (int a = 0; a < 3; a++) {
                const float ext = cell.max[a] - cell.min[a];
                const float scale = ext > 0.f ? (float)B / ext : 0.f;
                const float clo = std::fmax(tb.min[a], cell.min[a]), chi = std::fmin(tb.max[a], cell.max[a]);
                H[a][0][binOf(clo, cell.min[a], scale)]++;
                H[a][1][binOf(chi, cell.min[a], scale)]++;
            }
        }
        float best = intersectionCost * (float)n;
        for (int a = 0; a < 3; a++) {
            const float ext = cell.max[a] - cell.min[a];
            if (!(ext > 0.f)) continue;
            uint32_t nl = 0, ends = 0;
            for (int b = 1; b < B; b++) {
                nl += H[a][0][b - 1];
                ends += H[a][1][b - 1];
                const uint32_t nr = (uint32_t)n - ends;
                const float p = cell.min[a] + ext * ((float)b * (1.0f / (float)B));
                float l[3] = {d[0], d[1], d[2]}, r[3] = {d[0], d[1], d[2]};
                l[a] = p - cell.min[a];
                r[a] = cell.max[a] - p;
                const float pl = halfArea(l[0], l[1], l[2]) / sa, pr = halfArea(r[0], r[1], r[2]) / sa;
                float cost = traversalCost + intersectionCost * (pl * (float)nl + pr * (float)nr);
                if ((nl == 0 || nr == 0) && !(pl == 1.f || pr == 1.f)) cost *= 0.8f;
                if (cost < best) best = cost, bestA = a, bestP = p;
            }
        }
    }
    if (bestA < 0) return std::make_unique<KDTreeElement>(cell, std::move(refs));
    std::vector<uint32_t> L, R;
    for (uint32_t t : refs) {
        const BoundingBox &tb = triangles[t].boundingBox;
        const float clo = std::fmax(tb.min[bestA], cell.min[bestA]), chi = std::fmin(tb.max[bestA], cell.max[bestA]);
        if (clo < bestP || (clo == bestP && chi == bestP)) L.push_back(t);
        if (chi > bestP) R.push_back(t);
    }
    std::vector<uint32_t>().swap(refs);
    BoundingBox lc = cell, rc = cell;
    lc.max[bestA] = bestP;
    rc.min[bestA] = bestP;
    std::unique_ptr<KDTreeElement> l, r;
    if (L.size() + R.size() > 4096) {
#pragma omp task shared(l, L, lc)
        l = rec(std::move(L), lc, depth + 1);
#pragma omp task shared(r, R, rc)
        r = rec(std::move(R), rc, depth + 1);
#pragma omp taskwait
    } else {
        l = rec(std::move(L), lc, depth + 1);
        r = rec(std::move(R), rc, depth + 1);
    }
    auto node = std::make_unique<KDTreeElement>();
    node->boundingBox = cell;
    node->left = std::move(l);
    node->right = std::move(r);
    node->splitPlane.axis = (uint8_t)bestA;
    node->splitPlane.distance = bestP;
    node->log();
    return node;
}

}  // namespace pray
