// GPU tree build (no reference counterpart on the GPU: the reference builds
// its median and SAH KD trees on the CPU, kd_tree.cpp:25-84,
// kd_tree_sah.cpp:32-296).
//
// The tree is a bounding interval hierarchy over Morton-ordered triangles,
// the GPU counterpart of the reference's object-median tree: triangles are
// sorted by the 30-bit Morton code of their bounding-box centre in the scene
// box, and the binary radix tree over the sorted codes (Karras 2012,
// "Maximizing parallelism in the construction of BVHs, octrees, and k-d
// trees") gives every inner node in one parallel step.  An inner node splits
// where the highest differing code bit of its range flips, so that bit's axis
// is its split axis; equal codes are told apart by their sorted position.
// Subtrees of at most `leaf_size` triangles become leaves.
//
// The result is the traversal's device layout directly: inner node i (the
// radix tree's numbering, root 0) keeps its children pair in pair slot i+1,
// each child's clip plane being its subtree's bound along node i's axis (left
// child: upper bound, right child: lower bound), so the bounding-interval
// form of the node step applies unchanged.  Leaf runs are the sorted
// positions of their triangles; every triangle is in exactly one leaf.
//
// Kernels, in order: k_prep (edges, normals, boxes), k_box_partial /
// k_box_final (scene box), k_morton, a stable radix sort of (code, index),
// k_radix_tree, k_levels_init / k_levels_down (inner nodes in level order),
// k_levels_up (boxes, counts, collapsed heights, deepest level first),
// k_write_nodes and k_records.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <deque>
#include <memory>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "../../include/pray_hip.h"
#include "context.hpp"

namespace prb {

using prt::LEAF_TAG;
using prt::REF_LAST;

constexpr uint32_t MORTON_BITS = 30;
constexpr uint32_t LEAF_BIT = 0x80000000u;  // child is a single triangle (sorted position)
constexpr int BOX_BLOCK = 256;
constexpr unsigned BOX_PARTIALS = 1024;

struct Tri {
    const float *abc;   // n x 9
    float *a, *ab, *ac, *normal;  // n x 3 each (normal may be null)
    float *lo, *hi;     // triangle boxes, n x 3 each
};

// Edges, normal and box of one triangle, with the host precompute's
// operation order (ab = b - a, ac = c - a, normal = normalize(ab x ac), box
// corners by min / max).
static __global__ void k_prep(uint32_t n, Tri T) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *p = T.abc + 9 * (size_t)i;
    const float ax = p[0], ay = p[1], az = p[2], bx = p[3], by = p[4], bz = p[5], cx = p[6], cy = p[7], cz = p[8];
    const float abx = bx - ax, aby = by - ay, abz = bz - az;
    const float acx = cx - ax, acy = cy - ay, acz = cz - az;
    float *o = T.a + 3 * (size_t)i;
    o[0] = ax, o[1] = ay, o[2] = az;
    o = T.ab + 3 * (size_t)i;
    o[0] = abx, o[1] = aby, o[2] = abz;
    o = T.ac + 3 * (size_t)i;
    o[0] = acx, o[1] = acy, o[2] = acz;
    if (T.normal) {
        float nx = aby * acz - abz * acy, ny = abz * acx - abx * acz, nz = abx * acy - aby * acx;
        const float n2 = nx * nx + ny * ny + nz * nz;
        if (n2 > 0.f) {
            const float inv = 1.0f / sqrtf(n2);
            nx *= inv, ny *= inv, nz *= inv;
        }
        o = T.normal + 3 * (size_t)i;
        o[0] = nx, o[1] = ny, o[2] = nz;
    }
    // std::min(x, std::min(y, z)) of the host: min(y, z) = z < y ? z : y
    auto mn = [](float x, float y, float z) {
        const float m = z < y ? z : y;
        return m < x ? m : x;
    };
    auto mx = [](float x, float y, float z) {
        const float m = y < z ? z : y;
        return x < m ? m : x;
    };
    float *l = T.lo + 3 * (size_t)i, *h = T.hi + 3 * (size_t)i;
    l[0] = mn(ax, bx, cx), l[1] = mn(ay, by, cy), l[2] = mn(az, bz, cz);
    h[0] = mx(ax, bx, cx), h[1] = mx(ay, by, cy), h[2] = mx(az, bz, cz);
}

// Scene box: per-block minima and maxima, then one block over the partials.
static __global__ __launch_bounds__(BOX_BLOCK) void k_box_partial(uint32_t n, const float *lo, const float *hi,
                                                                  float *partial) {
    __shared__ float s[6][BOX_BLOCK];
    float v[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t i = blockIdx.x * BOX_BLOCK + threadIdx.x; i < n; i += gridDim.x * BOX_BLOCK)
        for (int k = 0; k < 3; k++) {
            v[k] = fminf(v[k], lo[3 * (size_t)i + k]);
            v[3 + k] = fmaxf(v[3 + k], hi[3 * (size_t)i + k]);
        }
    for (int k = 0; k < 6; k++) s[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = BOX_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 6; k++)
                s[k][threadIdx.x] = k < 3 ? fminf(s[k][threadIdx.x], s[k][threadIdx.x + w])
                                          : fmaxf(s[k][threadIdx.x], s[k][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

static __global__ __launch_bounds__(BOX_BLOCK) void k_box_final(uint32_t n_partials, const float *partial,
                                                                float *box) {
    __shared__ float s[6][BOX_BLOCK];
    float v[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t i = threadIdx.x; i < n_partials; i += BOX_BLOCK)
        for (int k = 0; k < 6; k++)
            v[k] = k < 3 ? fminf(v[k], partial[6 * i + k]) : fmaxf(v[k], partial[6 * i + k]);
    for (int k = 0; k < 6; k++) s[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = BOX_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 6; k++)
                s[k][threadIdx.x] = k < 3 ? fminf(s[k][threadIdx.x], s[k][threadIdx.x + w])
                                          : fmaxf(s[k][threadIdx.x], s[k][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x < 6) box[threadIdx.x] = s[threadIdx.x][0];
}

// Spreads the low 10 bits of v to every third bit.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Quantised box centre in [0, 1024) per axis: ((centre - lo) * inv) * 1024,
// clamped; x takes bits 3k+2, y 3k+1, z 3k.
static __global__ void k_morton(uint32_t n, const float *lo, const float *hi, const float *box, uint32_t *keys,
                                uint32_t *vals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
        const float ext = box[3 + k] - box[k];
        const float inv = ext > 0.f ? 1.0f / ext : 0.f;
        const float c = (lo[3 * (size_t)i + k] + hi[3 * (size_t)i + k]) * 0.5f;
        const float u = ((c - box[k]) * inv) * 1024.0f;
        q[k] = (uint32_t)fmaxf(0.f, fminf(1023.f, u));
    }
    keys[i] = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
    vals[i] = i;
}

// Length of the common prefix of (key, position) at sorted positions i and
// j, or -1 when j is outside [0, n).
__device__ __forceinline__ int delta(const uint32_t *keys, uint32_t n, int i, int j) {
    if (j < 0 || j >= (int)n) return -1;
    const uint32_t a = keys[i], b = keys[j];
    if (a != b) return __clz(a ^ b);
    return 32 + __clz((uint32_t)i ^ (uint32_t)j);
}

struct Radix {
    const uint32_t *keys;
    uint32_t n;
    uint32_t *left, *right;   // child: inner node index, or LEAF_BIT | sorted position
    uint32_t *first, *last;   // sorted range of each inner node
};

// One inner node per thread (Karras 2012, algorithm 2).
static __global__ void k_radix_tree(Radix R) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= (int)R.n - 1) return;
    const int d = delta(R.keys, R.n, i, i + 1) - delta(R.keys, R.n, i, i - 1) >= 0 ? 1 : -1;
    const int dmin = delta(R.keys, R.n, i, i - d);
    int lmax = 2;
    while (delta(R.keys, R.n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta(R.keys, R.n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(R.keys, R.n, i, j);
    int s = 0;
    for (int div = 2;; div *= 2) {
        const int t = (l + div - 1) / div;
        if (delta(R.keys, R.n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    const int g = i + s * d + min(d, 0);
    const int lo = min(i, j), hi = max(i, j);
    const uint32_t lc = lo == g ? (LEAF_BIT | (uint32_t)g) : (uint32_t)g;
    const uint32_t rc = hi == g + 1 ? (LEAF_BIT | (uint32_t)(g + 1)) : (uint32_t)(g + 1);
    R.left[i] = lc, R.right[i] = rc;
    R.first[i] = (uint32_t)lo, R.last[i] = (uint32_t)hi;
}

struct Up {
    uint32_t n, leaf_size;
    const uint32_t *vals;       // sorted position -> triangle
    const float *tlo, *thi;     // triangle boxes
    const uint32_t *left, *right;
    float *nlo, *nhi;           // inner node boxes, (n-1) x 3
    uint32_t *count, *height;   // triangles under an inner node; height with collapsed leaves
};

__device__ __forceinline__ void child_info(const Up &U, uint32_t c, float lo[3], float hi[3], uint32_t &cnt,
                                           uint32_t &h) {
    if (c & LEAF_BIT) {
        const uint32_t t = U.vals[c & ~LEAF_BIT];
        for (int k = 0; k < 3; k++) lo[k] = U.tlo[3 * (size_t)t + k], hi[k] = U.thi[3 * (size_t)t + k];
        cnt = 1, h = 0;
    } else {
        for (int k = 0; k < 3; k++) lo[k] = U.nlo[3 * (size_t)c + k], hi[k] = U.nhi[3 * (size_t)c + k];
        cnt = U.count[c];
        h = cnt <= U.leaf_size ? 0u : U.height[c];
    }
}

// The inner nodes in level order, so that boxes are combined one level per
// launch, deepest first: a bottom-up walk in one launch would need a
// device-scope fence per step, and on a part whose L2s are not coherent with
// each other each such fence writes back and invalidates the L2.  Level d+1
// lists the inner children of level d's nodes.  The radix tree is less than
// 65 levels deep: each level lengthens the common prefix of the 64-bit
// (key, position) pairs.
constexpr uint32_t UP_LEVELS = 72;

struct Levels {
    const uint32_t *left, *right;
    uint32_t n_inner;
    uint32_t *order;  // inner nodes, level after level (n_inner entries)
    uint32_t *count;  // nodes per level (UP_LEVELS + 1)
    uint32_t *start;  // first entry of each level in order (UP_LEVELS + 1)
};

static __global__ void k_levels_init(Levels L) {
    for (uint32_t j = threadIdx.x; j <= UP_LEVELS; j += blockDim.x) {
        L.count[j] = j == 0 ? 1u : 0u;
        L.start[j] = 0u;
    }
    if (threadIdx.x == 0) L.order[0] = 0u;  // the root
}

// Appends the inner children of level d's nodes to level d+1, in any order
// (a node's values depend on its children only); one atomic per wave.
static __global__ __launch_bounds__(256) void k_levels_down(Levels L, uint32_t d) {
    const uint32_t s = L.start[d], cnt = L.count[d], s1 = s + cnt;
    if (blockIdx.x == 0 && threadIdx.x == 0) L.start[d + 1] = s1;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < cnt; i0 += stride) {  // wave-uniform
        const uint32_t i = i0 + lane;
        uint32_t l = LEAF_BIT, r = LEAF_BIT;
        if (i < cnt && s + i < L.n_inner) {
            const uint32_t node = L.order[s + i];
            l = L.left[node], r = L.right[node];
        }
        const bool hl = !(l & LEAF_BIT), hr = !(r & LEAF_BIT);
        const uint64_t bl = __ballot(hl), br = __ballot(hr);
        const uint32_t off = (uint32_t)(__popcll(bl & below) + __popcll(br & below));
        const uint32_t total = (uint32_t)(__popcll(bl) + __popcll(br));
        uint32_t base = 0u;
        if (lane == 0 && total) base = atomicAdd(&L.count[d + 1], total);
        base = (uint32_t)__shfl((int)base, 0);
        const uint32_t j = s1 + base + off;
        if (hl && j < L.n_inner) L.order[j] = l;
        if (hr && j + (hl ? 1u : 0u) < L.n_inner) L.order[j + (hl ? 1u : 0u)] = r;
    }
}

// Boxes, counts and collapsed heights of one level's nodes (their inner
// children are on the next level, combined by the previous launch).
static __global__ void k_levels_up(Up U, const uint32_t *order, uint32_t s, uint32_t cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    const uint32_t node = order[s + i];
    float llo[3], lhi[3], rlo[3], rhi[3];
    uint32_t lc, lh, rc, rh;
    child_info(U, U.left[node], llo, lhi, lc, lh);
    child_info(U, U.right[node], rlo, rhi, rc, rh);
    for (int q = 0; q < 3; q++) {
        U.nlo[3 * (size_t)node + q] = fminf(llo[q], rlo[q]);
        U.nhi[3 * (size_t)node + q] = fmaxf(lhi[q], rhi[q]);
    }
    const uint32_t c = lc + rc;
    U.count[node] = c;
    U.height[node] = c <= U.leaf_size ? 0u : 1u + max(lh, rh);
}

struct Writer {
    uint32_t n, leaf_size;
    const uint32_t *keys;
    const uint32_t *left, *right, *first, *last, *count;
    const float *nlo, *nhi;
    uint2 *nodes;          // 2n entries (pairs 0 .. n-1)
    uint8_t *last_flag;    // sorted positions that end a leaf
};

// Split axis of inner node c: the axis of the highest code bit that differs
// over its range, or its box's longest axis when all codes are equal.
__device__ __forceinline__ uint32_t node_axis(const Writer &W, uint32_t c) {
    const uint32_t x = W.keys[W.first[c]] ^ W.keys[W.last[c]];
    if (x) {
        const uint32_t bit = 31u - (uint32_t)__clz(x);
        const uint32_t m = bit % 3u;
        return m == 2u ? 0u : (m == 1u ? 1u : 2u);
    }
    const float ex = W.nhi[3 * (size_t)c] - W.nlo[3 * (size_t)c];
    const float ey = W.nhi[3 * (size_t)c + 1] - W.nlo[3 * (size_t)c + 1];
    const float ez = W.nhi[3 * (size_t)c + 2] - W.nlo[3 * (size_t)c + 2];
    return (ex >= ey && ex >= ez) ? 0u : (ey >= ez ? 1u : 2u);
}

__device__ __forceinline__ bool is_leaf(const Writer &W, uint32_t c) {
    return (c & LEAF_BIT) || W.count[c] <= W.leaf_size;
}

__device__ __forceinline__ uint32_t child_word(const Writer &W, uint32_t c) {
    if (c & LEAF_BIT) {
        W.last_flag[c & ~LEAF_BIT] = 1;
        return ((c & ~LEAF_BIT) << 2) | LEAF_TAG;
    }
    if (W.count[c] <= W.leaf_size) {
        W.last_flag[W.last[c]] = 1;
        return (W.first[c] << 2) | LEAF_TAG;
    }
    return ((2u * (c + 1u)) << 2) | node_axis(W, c);
}

// Bound of child c along `axis`: its upper bound for a left child, lower
// bound for a right child.
__device__ __forceinline__ float child_bound(const Writer &W, const Up &U, uint32_t c, uint32_t axis, bool upper) {
    if (c & LEAF_BIT) {
        const uint32_t t = U.vals[c & ~LEAF_BIT];
        return upper ? U.thi[3 * (size_t)t + axis] : U.tlo[3 * (size_t)t + axis];
    }
    return upper ? W.nhi[3 * (size_t)c + axis] : W.nlo[3 * (size_t)c + axis];
}

static __global__ void k_write_nodes(Writer W, Up U) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {  // the root's entry: node 1
        const uint32_t w = (W.n < 2 || W.count[0] <= W.leaf_size) ? LEAF_TAG : (((2u << 2)) | node_axis(W, 0));
        if (w == LEAF_TAG) W.last_flag[W.n - 1] = 1;
        W.nodes[1] = make_uint2(0u, w);
    }
    if (W.n < 2 || i >= W.n - 1 || W.count[i] <= W.leaf_size) return;
    const uint32_t axis = node_axis(W, i);
    const uint32_t l = W.left[i], r = W.right[i];
    W.nodes[2 * (size_t)(i + 1)] = make_uint2(__float_as_uint(child_bound(W, U, l, axis, true)), child_word(W, l));
    W.nodes[2 * (size_t)(i + 1) + 1] =
        make_uint2(__float_as_uint(child_bound(W, U, r, axis, false)), child_word(W, r));
}

// Leaf references in sorted order (triangle | last flag).
static __global__ void k_morton_refs(uint32_t n, const uint32_t *vals, const uint8_t *last_flag, uint32_t *refs) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    refs[k] = vals[k] | (last_flag[k] ? REF_LAST : 0u);
}

// ---------------------------------------------------------------------------
// Binned-SAH KD tree, built breadth first: one pass of kernels per tree level
// over all triangle references of that level (references of one node are
// contiguous, in increasing triangle order).
//   k_sah_small     per node of at most SAH_SMALL references: one wave bins
//                   them and picks the plane itself, with no histogram in
//                   global memory (the same counts, costs and choice as the
//                   two kernels below); for each larger node it clears that
//                   node's histogram
//   k_sah_bins      per larger node and axis, counts of clipped reference
//                   boxes starting and ending in each of SAH_BINS (64) bins
//   k_sah_split     per larger node: the cheapest plane on a bin boundary against
//                   the leaf cost (the reference's SAH cost with its empty
//                   -side factor 0.8, kd_tree_sah.cpp:141-166), or a leaf
//   k_sah_classify  per reference: left if its clipped box starts below the
//                   plane (or lies in it), right if it ends above it
//   k_sah_counts    per node: child counts from the scanned flags
//   k_sah_emit      per node: its own entry in the device layout (leaf run,
//                   empty leaf, or children pair in BFS order) and the next
//                   level's nodes
//   k_sah_scatter   per reference: stable partition into the next level, or
//                   into the final leaf runs
// Both children of a split share the plane as their clip (the KD form of the
// node step).  Everything is deterministic; the oracle's twin reproduces it
// word for word.
constexpr int SAH_BINS = 64;
constexpr int SAH_HIST = 3 * 2 * SAH_BINS;  // per node: axis x {start, end} x bin
constexpr int SAH_CHUNK = 1024;             // references per block of k_sah_bins
constexpr int SAH_SLOTS = 16;               // LDS histograms per block of k_sah_bins
constexpr uint32_t SAH_SMALL = 64;          // nodes up to this many references: k_sah_small

struct LNode {
    float lo[3], hi[3];     // cell
    uint32_t begin, count;  // references
    uint32_t entry;         // index of this node's entry in the device layout
    float plane;            // clip plane written with the entry (the parent's split)
};

struct SahParams {
    uint32_t leaf_size, max_depth;
    float ct, ci;
};

__device__ __forceinline__ int sah_bin(float x, float lo, float scale) {
    return (int)fmaxf(0.f, fminf((float)(SAH_BINS - 1), floorf((x - lo) * scale)));
}

__device__ __forceinline__ float half_area(float dx, float dy, float dz) { return dx * dy + dy * dz + dz * dx; }

// A block bins SAH_CHUNK consecutive references (sorted by node) into LDS
// histograms for the first SAH_SLOTS nodes of its range, and the rest
// straight into global memory; then it adds the nonzero LDS counts to the
// global histograms (a node may span several blocks).
static __global__ __launch_bounds__(256) void k_sah_bins(uint32_t n_refs, const uint32_t *refs,
                                                         const uint32_t *ref_node, const float *tlo,
                                                         const float *thi, const LNode *nodes, uint32_t *hist) {
    __shared__ uint32_t h[SAH_SLOTS * SAH_HIST];
    const uint32_t first = blockIdx.x * SAH_CHUNK;
    if (first >= n_refs) return;
    const uint32_t last = min(first + (uint32_t)SAH_CHUNK, n_refs) - 1;
    const uint32_t k0 = ref_node[first];
    const uint32_t nslots = min(ref_node[last] - k0 + 1, (uint32_t)SAH_SLOTS);
    for (uint32_t j = threadIdx.x; j < nslots * SAH_HIST; j += 256) h[j] = 0u;
    __syncthreads();
    for (uint32_t i = first + threadIdx.x; i <= last; i += 256) {
        const uint32_t k = ref_node[i], t = refs[i];
        const LNode &N = nodes[k];
        if (N.count <= SAH_SMALL) continue;  // binned by k_sah_small
        const uint32_t slot = k - k0;
        for (int a = 0; a < 3; a++) {
            const float ext = N.hi[a] - N.lo[a];
            const float scale = ext > 0.f ? (float)SAH_BINS / ext : 0.f;
            const float clo = fmaxf(tlo[3 * (size_t)t + a], N.lo[a]);
            const float chi = fminf(thi[3 * (size_t)t + a], N.hi[a]);
            const int bs = a * 2 * SAH_BINS + sah_bin(clo, N.lo[a], scale);
            const int be = a * 2 * SAH_BINS + SAH_BINS + sah_bin(chi, N.lo[a], scale);
            if (slot < (uint32_t)SAH_SLOTS) {
                atomicAdd(&h[slot * SAH_HIST + bs], 1u);
                atomicAdd(&h[slot * SAH_HIST + be], 1u);
            } else {
                atomicAdd(&hist[(size_t)k * SAH_HIST + bs], 1u);
                atomicAdd(&hist[(size_t)k * SAH_HIST + be], 1u);
            }
        }
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < nslots * SAH_HIST; j += 256) {
        const uint32_t v = h[j];
        if (v) atomicAdd(&hist[(size_t)(k0 + j / SAH_HIST) * SAH_HIST + j % SAH_HIST], v);
    }
}

// One wave per node of at most SAH_SMALL references (blocks of 4 waves):
// lane i bins reference i, lane b (1..63) counts the references starting /
// ending below bin boundary b from the wave's bins and evaluates the plane at
// b on each axis with k_sah_split's arithmetic; the first minimum in
// (axis, bin) order below the leaf cost wins, as in k_sah_split's scan.
static __global__ __launch_bounds__(256) void k_sah_small(uint32_t n_nodes, const LNode *nodes, const uint32_t *refs,
                                                          const float *tlo, const float *thi, SahParams P,
                                                          uint32_t level, int *axis, float *pos, uint32_t *hist) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t k = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (k >= n_nodes) return;  // wave-uniform
    const LNode N = nodes[k];
    if (N.count > SAH_SMALL) {  // wave-uniform: clears its histogram for k_sah_bins / k_sah_split
        for (uint32_t j = lane; j < (uint32_t)SAH_HIST; j += 64) hist[(size_t)k * SAH_HIST + j] = 0u;
        return;
    }
    const float dx = N.hi[0] - N.lo[0], dy = N.hi[1] - N.lo[1], dz = N.hi[2] - N.lo[2];
    const float sa = half_area(dx, dy, dz);
    if (!(N.count > P.leaf_size && level < P.max_depth && sa > 0.f)) {
        if (lane == 0) axis[k] = -1, pos[k] = 0.f;
        return;
    }
    // bins of reference `lane`: start bins in bytes 0..2, end bins in 0..2 of e
    uint32_t s = 0, e = 0;
    float scale[3];
    for (int a = 0; a < 3; a++) {
        const float ext = N.hi[a] - N.lo[a];
        scale[a] = ext > 0.f ? (float)SAH_BINS / ext : 0.f;
    }
    if (lane < N.count) {
        const uint32_t t = refs[N.begin + lane];
        for (int a = 0; a < 3; a++) {
            const float clo = fmaxf(tlo[3 * (size_t)t + a], N.lo[a]);
            const float chi = fminf(thi[3 * (size_t)t + a], N.hi[a]);
            s |= (uint32_t)sah_bin(clo, N.lo[a], scale[a]) << (8 * a);
            e |= (uint32_t)sah_bin(chi, N.lo[a], scale[a]) << (8 * a);
        }
    }
    uint32_t nl[3] = {0, 0, 0}, ne[3] = {0, 0, 0};  // references starting / ending below boundary `lane`
    for (uint32_t i = 0; i < N.count; i++) {
        const uint32_t si = (uint32_t)__shfl((int)s, (int)i), ei = (uint32_t)__shfl((int)e, (int)i);
        for (int a = 0; a < 3; a++) {
            nl[a] += ((si >> (8 * a)) & 0xFFu) < lane ? 1u : 0u;
            ne[a] += ((ei >> (8 * a)) & 0xFFu) < lane ? 1u : 0u;
        }
    }
    // this lane's best plane: (cost, index a * 64 + b); a later axis only
    // replaces an earlier one when strictly cheaper
    float best = INFINITY;
    uint32_t best_i = 0xFFFFFFFFu;
    if (lane >= 1) {
        for (int a = 0; a < 3; a++) {
            const float ext = N.hi[a] - N.lo[a];
            if (!(ext > 0.f)) continue;
            const uint32_t n_left = nl[a], n_right = N.count - ne[a];
            const float p = N.lo[a] + ext * ((float)lane * (1.0f / (float)SAH_BINS));
            float l[3] = {dx, dy, dz}, r[3] = {dx, dy, dz};
            l[a] = p - N.lo[a];
            r[a] = N.hi[a] - p;
            const float pl = half_area(l[0], l[1], l[2]) / sa, pr = half_area(r[0], r[1], r[2]) / sa;
            float cost = P.ct + P.ci * (pl * (float)n_left + pr * (float)n_right);
            if ((n_left == 0 || n_right == 0) && !(pl == 1.f || pr == 1.f)) cost *= 0.8f;
            if (cost < best) best = cost, best_i = (uint32_t)a * SAH_BINS + lane;
        }
    }
    // wave minimum of (cost, index): the first minimum of the sequential scan
    for (int off = 32; off > 0; off >>= 1) {
        const float oc = __shfl_xor(best, off);
        const uint32_t oi = (uint32_t)__shfl_xor((int)best_i, off);
        if (oc < best || (oc == best && oi < best_i)) best = oc, best_i = oi;
    }
    if (lane == 0) {
        const float leaf_cost = P.ci * (float)N.count;
        if (best < leaf_cost) {
            const int a = (int)(best_i / SAH_BINS);
            const uint32_t b = best_i % SAH_BINS;
            axis[k] = a;
            pos[k] = N.lo[a] + (N.hi[a] - N.lo[a]) * ((float)b * (1.0f / (float)SAH_BINS));
        } else {
            axis[k] = -1;
            pos[k] = 0.f;
        }
    }
}

// axis[k] = split axis or -1 for a leaf; pos[k] = split plane.
static __global__ void k_sah_split(uint32_t n_nodes, const LNode *nodes, const uint32_t *hist, SahParams P,
                                   uint32_t level, int *axis, float *pos) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_nodes) return;
    const LNode N = nodes[k];
    if (N.count <= SAH_SMALL) return;  // decided by k_sah_small
    int best_a = -1;
    float best_p = 0.f;
    const float dx = N.hi[0] - N.lo[0], dy = N.hi[1] - N.lo[1], dz = N.hi[2] - N.lo[2];
    const float sa = half_area(dx, dy, dz);
    if (N.count > P.leaf_size && level < P.max_depth && sa > 0.f) {
        const float leaf_cost = P.ci * (float)N.count;
        float best = leaf_cost;
        const uint32_t *H = hist + (size_t)k * SAH_HIST;
        for (int a = 0; a < 3; a++) {
            const float ext = N.hi[a] - N.lo[a];
            if (!(ext > 0.f)) continue;
            uint32_t n_left = 0, ends = 0;
            for (int b = 1; b < SAH_BINS; b++) {
                n_left += H[a * 2 * SAH_BINS + b - 1];
                ends += H[a * 2 * SAH_BINS + SAH_BINS + b - 1];
                const uint32_t n_right = N.count - ends;
                const float p = N.lo[a] + ext * ((float)b * (1.0f / (float)SAH_BINS));
                float l[3] = {dx, dy, dz}, r[3] = {dx, dy, dz};
                l[a] = p - N.lo[a];
                r[a] = N.hi[a] - p;
                const float pl = half_area(l[0], l[1], l[2]) / sa, pr = half_area(r[0], r[1], r[2]) / sa;
                float cost = P.ct + P.ci * (pl * (float)n_left + pr * (float)n_right);
                if ((n_left == 0 || n_right == 0) && !(pl == 1.f || pr == 1.f)) cost *= 0.8f;
                if (cost < best) best = cost, best_a = a, best_p = p;
            }
        }
    }
    axis[k] = best_a;
    pos[k] = best_p;
}

// f[i] = (reference i goes left, goes right), (0, 0) for references of
// leaves; one extra zero entry at n_refs for the scan.
static __global__ void k_sah_classify(uint32_t n_refs, const uint32_t *refs, const uint32_t *ref_node,
                                      const float *tlo, const float *thi, const LNode *nodes, const int *axis,
                                      const float *pos, uint2 *f) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == n_refs) f[i] = make_uint2(0u, 0u);
    if (i >= n_refs) return;
    const uint32_t k = ref_node[i], t = refs[i];
    const int a = axis[k];
    if (a < 0) {
        f[i] = make_uint2(0u, 0u);
        return;
    }
    const float p = pos[k];
    const float clo = fmaxf(tlo[3 * (size_t)t + a], nodes[k].lo[a]);
    const float chi = fminf(thi[3 * (size_t)t + a], nodes[k].hi[a]);
    f[i] = make_uint2((clo < p || (clo == p && chi == p)) ? 1u : 0u, chi > p ? 1u : 0u);
}

// Per-node values for the node scan, from the scanned reference flags s:
// (splits (1 per split node), references of the children, references of
// leaves, 0); one extra zero entry at n_nodes.
static __global__ void k_sah_counts(uint32_t n_nodes, const LNode *nodes, const int *axis, const uint2 *s,
                                    uint4 *v) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == n_nodes) v[k] = make_uint4(0u, 0u, 0u, 0u);
    if (k >= n_nodes) return;
    const LNode N = nodes[k];
    if (axis[k] < 0) {
        v[k] = make_uint4(0u, 0u, N.count, 0u);
    } else {
        const uint2 b = s[N.begin], e = s[N.begin + N.count];
        v[k] = make_uint4(1u, (e.x - b.x) + (e.y - b.y), 0u, 0u);
    }
}

struct Emit {
    uint32_t n_nodes;
    const LNode *nodes;
    const int *axis;
    const float *pos;
    const uint2 *s;   // scanned reference flags (left, right)
    const uint4 *sn;  // scanned node values (splits, child references, leaf references, 0)
    uint32_t pair_base, leaf_base;
    uint2 *out;          // device layout entries
    LNode *next;         // next level's nodes
};

static __global__ void k_sah_emit(Emit E) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= E.n_nodes) return;
    const LNode N = E.nodes[k];
    const int a = E.axis[k];
    const uint4 sn = E.sn[k];
    if (a < 0) {
        const uint32_t w = N.count ? (((E.leaf_base + sn.z) << 2) | LEAF_TAG) : ((prt::LEAF_EMPTY << 2) | LEAF_TAG);
        E.out[N.entry] = make_uint2(__float_as_uint(N.plane), w);
        return;
    }
    const uint32_t q = E.pair_base + sn.x;
    E.out[N.entry] = make_uint2(__float_as_uint(N.plane), ((2u * q) << 2) | (uint32_t)a);
    const uint2 b = E.s[N.begin], e = E.s[N.begin + N.count];
    const uint32_t nl = e.x - b.x;
    const float p = E.pos[k];
    LNode L = N, R = N;
    L.hi[a] = p, R.lo[a] = p;
    L.begin = sn.y, L.count = nl, L.entry = 2 * q, L.plane = p;
    R.begin = sn.y + nl;
    R.count = e.y - b.y;
    R.entry = 2 * q + 1, R.plane = p;
    E.next[2 * sn.x] = L;
    E.next[2 * sn.x + 1] = R;
}

struct Scatter {
    uint32_t n_refs;
    const uint32_t *refs, *ref_node;
    const LNode *nodes;
    const int *axis;
    const uint2 *f, *s;  // reference flags (left, right) and their scan
    const uint4 *sn;     // scanned node values, as in Emit
    uint32_t leaf_base;
    uint32_t *next_refs, *next_node;  // next level
    uint32_t *leaf_refs;              // final leaf runs (triangle | last flag)
};

static __global__ void k_sah_scatter(Scatter S) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.n_refs) return;
    const uint32_t k = S.ref_node[i], t = S.refs[i];
    const LNode N = S.nodes[k];
    const uint4 sn = S.sn[k];
    if (S.axis[k] < 0) {
        const uint32_t r = i - N.begin;
        S.leaf_refs[S.leaf_base + sn.z + r] = t | (r + 1 == N.count ? REF_LAST : 0u);
        return;
    }
    const uint2 b = S.s[N.begin], e = S.s[N.begin + N.count], f = S.f[i], si = S.s[i];
    const uint32_t nl = e.x - b.x;
    if (f.x) {
        const uint32_t j = sn.y + (si.x - b.x);
        S.next_refs[j] = t;
        S.next_node[j] = 2 * sn.x;
    }
    if (f.y) {
        const uint32_t j = sn.y + nl + (si.y - b.y);
        S.next_refs[j] = t;
        S.next_node[j] = 2 * sn.x + 1;
    }
}

// Packed records of the final leaf runs.
static __global__ void k_pack_leaf_records(uint32_t n_refs, const uint32_t *refs, const float *a, const float *ab,
                                           const float *ac, float4 *recs) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_refs) return;
    const uint32_t ref = refs[k];
    const size_t o = 3 * (size_t)(ref & ~REF_LAST);
    recs[3 * (size_t)k] = make_float4(a[o], a[o + 1], a[o + 2], __uint_as_float(ref));
    recs[3 * (size_t)k + 1] = make_float4(ab[o], ab[o + 1], ab[o + 2], 0.f);
    recs[3 * (size_t)k + 2] = make_float4(ac[o], ac[o + 1], ac[o + 2], 0.f);
}

static __global__ void k_sah_root(const float *box, uint32_t n, LNode *nodes, uint32_t *refs, uint32_t *ref_node) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        LNode r;
        for (int k = 0; k < 3; k++) r.lo[k] = box[k], r.hi[k] = box[3 + k];
        r.begin = 0, r.count = n, r.entry = 1, r.plane = 0.f;
        nodes[0] = r;
    }
    if (i < n) refs[i] = i, ref_node[i] = 0u;
}

}  // namespace prb

namespace prc {
namespace {

// A device buffer that grows (by at least half) and optionally keeps its
// contents; freed on every exit.
struct Grow {
    void *p = nullptr;
    size_t cap = 0;
    Grow() = default;
    Grow(const Grow &) = delete;  // owns p
    Grow &operator=(const Grow &) = delete;
    ~Grow() {
        if (p) (void)hipFree(p);
    }
    int reserve(pray_ctx *c, size_t bytes, bool keep) {
        if (bytes <= cap) return PRAY_OK;
        const size_t nc = std::max(bytes, 2 * cap);
        void *q = nullptr;
        HIPC(c, hipMalloc(&q, nc));
        if (keep && p && cap) HIPC(c, hipMemcpyAsync(q, p, cap, hipMemcpyDeviceToDevice, c->stream));
        if (p) {
            HIPC(c, hipStreamSynchronize(c->stream));
            HIPC(c, hipFree(p));
        }
        p = q, cap = nc;
        return PRAY_OK;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Numbered device buffers of one call site, taken in the same order on every
// call.
struct Pool {
    std::deque<Grow> slots;  // a deque: taking a new slot moves no other
    size_t next = 0;
    template <class T> int get(pray_ctx *c, T **p, size_t count) {
        if (next == slots.size()) slots.emplace_back();
        Grow &g = slots[next++];
        const int rc = g.reserve(c, std::max<size_t>(count, 1) * sizeof(T), false);
        if (rc) return rc;
        *p = g.as<T>();
        return PRAY_OK;
    }
};

// The tree builds' device buffers, kept by the context from one build to the
// next (a scene rebuilt every frame allocates nothing after its first build)
// and freed by pray_destroy.  Builds on one context run one at a time.
struct BuildScratch {
    Pool prep, morton;
    Grow nodes[2], refs[2], rnode[2], f, s, axis, pos, hist, v, sn, tmp;  // the SAH levels
    Grow entries, leaf_refs;  // a build's output before it is copied into the scene
};

BuildScratch &build_scratch(pray_ctx *c) {
    if (!c->build_scratch) c->build_scratch = std::make_shared<BuildScratch>();
    return *static_cast<BuildScratch *>(c->build_scratch.get());
}

struct Built {
    uint32_t n_entries = 0, n_refs = 0, depth = 0;
};

// Morton-ordered bounding interval hierarchy: entries = 2n node entries,
// refs = n leaf references.
int build_morton(pray_ctx *c, hipStream_t st, uint32_t n, uint32_t leaf_size, const float *tlo, const float *thi,
                 const float *box, Grow &entries, Grow &refs, Built &out) {
    Pool &S = build_scratch(c).morton;
    S.next = 0;
    int rc = PRAY_OK;
    uint32_t *keys = nullptr, *vals = nullptr, *keys2 = nullptr, *vals2 = nullptr, *left = nullptr,
             *right = nullptr, *first = nullptr, *last = nullptr, *order = nullptr, *lcount = nullptr,
             *lstart = nullptr, *count = nullptr, *height = nullptr;
    float *nlo = nullptr, *nhi = nullptr;
    uint8_t *last_flag = nullptr;
    const size_t inner = n > 1 ? n - 1 : 1;
    if ((rc = S.get(c, &keys, n)) || (rc = S.get(c, &vals, n)) || (rc = S.get(c, &keys2, n)) ||
        (rc = S.get(c, &vals2, n)) || (rc = S.get(c, &left, inner)) || (rc = S.get(c, &right, inner)) ||
        (rc = S.get(c, &first, inner)) || (rc = S.get(c, &last, inner)) || (rc = S.get(c, &order, inner)) ||
        (rc = S.get(c, &lcount, prb::UP_LEVELS + 1)) || (rc = S.get(c, &lstart, prb::UP_LEVELS + 1)) ||
        (rc = S.get(c, &count, inner)) ||
        (rc = S.get(c, &height, inner)) || (rc = S.get(c, &nlo, 3 * inner)) || (rc = S.get(c, &nhi, 3 * inner)) ||
        (rc = S.get(c, &last_flag, n)))
        return rc;
    if ((rc = entries.reserve(c, 2 * (size_t)n * 8, false)) || (rc = refs.reserve(c, (size_t)n * 4, false)))
        return rc;
    uint2 *nodes = entries.as<uint2>();
    HIPC(c, hipMemsetAsync(nodes, 0, 2 * (size_t)n * 8, st));
    HIPC(c, hipMemsetAsync(last_flag, 0, (size_t)n, st));
    size_t sort_bytes = 0;
    HIPC(c, rocprim::radix_sort_pairs(nullptr, sort_bytes, keys, keys2, vals, vals2, n, 0u, prb::MORTON_BITS, st));
    uint8_t *sort_tmp = nullptr;
    if ((rc = S.get(c, &sort_tmp, sort_bytes))) return rc;
    const unsigned g = nblocks(n, 256);
    hipLaunchKernelGGL(prb::k_morton, dim3(g), dim3(256), 0, st, n, tlo, thi, box, keys, vals);
    HIPC(c, hipGetLastError());
    HIPC(c, rocprim::radix_sort_pairs(sort_tmp, sort_bytes, keys, keys2, vals, vals2, n, 0u, prb::MORTON_BITS, st));
    prb::Radix R{keys2, n, left, right, first, last};
    prb::Up U{n, leaf_size, vals2, tlo, thi, left, right, nlo, nhi, count, height};
    prb::Writer W{n, leaf_size, keys2, left, right, first, last, count, nlo, nhi, nodes, last_flag};
    if (n > 1) {
        hipLaunchKernelGGL(prb::k_radix_tree, dim3(nblocks(n - 1, 256)), dim3(256), 0, st, R);
        // inner nodes in level order (one launch per possible level: the
        // sizes are on the device), then one launch per level from the deepest
        const prb::Levels L{left, right, n - 1, order, lcount, lstart};
        hipLaunchKernelGGL(prb::k_levels_init, dim3(1), dim3(128), 0, st, L);
        const unsigned down_blocks = std::min<unsigned>(nblocks(n - 1, 256), 1024u);
        for (uint32_t d = 0; d < prb::UP_LEVELS; d++)
            hipLaunchKernelGGL(prb::k_levels_down, dim3(down_blocks), dim3(256), 0, st, L, d);
        HIPC(c, hipGetLastError());
        uint32_t lc[prb::UP_LEVELS + 1];
        HIPC(c, hipMemcpyAsync(lc, lcount, sizeof lc, hipMemcpyDeviceToHost, st));
        HIPC(c, hipStreamSynchronize(st));
        uint32_t ls[prb::UP_LEVELS + 1], total = 0, levels = 0;
        for (uint32_t d = 0; d <= prb::UP_LEVELS; d++) {
            ls[d] = total, total += lc[d];
            if (lc[d]) levels = d + 1;
        }
        if (lc[prb::UP_LEVELS] || total != n - 1)
            return fail(c, PRAY_ERR_UNSUPPORTED, "Morton build: %u inner nodes in %u levels (expected %u)", total,
                        levels, n - 1);
        for (uint32_t d = levels; d-- > 0;)
            hipLaunchKernelGGL(prb::k_levels_up, dim3(nblocks(lc[d], 256)), dim3(256), 0, st, U,
                               (const uint32_t *)order, ls[d], lc[d]);
    }
    hipLaunchKernelGGL(prb::k_write_nodes, dim3(nblocks(inner, 256)), dim3(256), 0, st, W, U);
    hipLaunchKernelGGL(prb::k_morton_refs, dim3(g), dim3(256), 0, st, n, (const uint32_t *)vals2,
                       (const uint8_t *)last_flag, refs.as<uint32_t>());
    HIPC(c, hipGetLastError());
    uint32_t depth = 0;
    if (n > 1) HIPC(c, hipMemcpyAsync(&depth, height, 4, hipMemcpyDeviceToHost, st));
    HIPC(c, hipStreamSynchronize(st));
    out.n_entries = 2 * n, out.n_refs = n, out.depth = depth;
    return PRAY_OK;
}

// Lane-wise sums of the scanned vectors (each lane is its own count).
struct AddU2 {
    __host__ __device__ uint2 operator()(const uint2 &a, const uint2 &b) const {
        return make_uint2(a.x + b.x, a.y + b.y);
    }
};
struct AddU4 {
    __host__ __device__ uint4 operator()(const uint4 &a, const uint4 &b) const {
        return make_uint4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
    }
};

// Exclusive lane-wise sum of in[0 .. count) into out with rocPRIM's device
// scan: one scan carries several counts at once.
template <class T, class Op>
int scan_sum(pray_ctx *c, hipStream_t st, Grow &tmp, const T *in, T *out, size_t count, T zero, Op op) {
    size_t bytes = 0;
    HIPC(c, rocprim::exclusive_scan(nullptr, bytes, in, out, zero, count, op, st));
    int rc = tmp.reserve(c, std::max<size_t>(bytes, 1), false);
    if (rc) return rc;
    HIPC(c, rocprim::exclusive_scan(tmp.p, bytes, in, out, zero, count, op, st));
    return PRAY_OK;
}

// Binned-SAH KD tree, one level at a time (kernels in prb).
int build_sah(pray_ctx *c, hipStream_t st, uint32_t n, const prb::SahParams &P, const float *tlo, const float *thi,
              const float *box, Grow &entries, Grow &leaf_refs, Built &out) {
    using prb::LNode;
    // f / s: reference flags (left, right) and their scan; v / sn: node values
    // (splits, child references, leaf references, 0) and their scan
    BuildScratch &X = build_scratch(c);
    Grow *nodes = X.nodes, *refs = X.refs, *rnode = X.rnode;
    Grow &f = X.f, &s = X.s, &axis = X.axis, &pos = X.pos, &hist = X.hist, &v = X.v, &sn = X.sn, &tmp = X.tmp;
    int rc = PRAY_OK;
    int cur = 0;
    // Up-front sizes from the triangle count (grown by doubling when a level
    // needs more): a level's nodes, its references with some duplication,
    // and the outputs.  Growing from small sizes reallocates (with a stream
    // synchronisation) at almost every level.
    const size_t est_nodes = std::max<size_t>(1024, (size_t)n / 2 + 1);
    const size_t est_refs = (size_t)n + (size_t)n / 4 + 1;
    for (Grow *g : {&nodes[0], &nodes[1]})
        if ((rc = g->reserve(c, est_nodes * sizeof(LNode), false))) return rc;
    for (Grow *g : {&refs[0], &refs[1], &rnode[0], &rnode[1]})
        if ((rc = g->reserve(c, est_refs * 4, false))) return rc;
    for (Grow *g : {&f, &s})
        if ((rc = g->reserve(c, (est_refs + 1) * sizeof(uint2), false))) return rc;
    for (Grow *g : {&axis, &pos})
        if ((rc = g->reserve(c, (est_nodes + 1) * 4, false))) return rc;
    for (Grow *g : {&v, &sn})
        if ((rc = g->reserve(c, (est_nodes + 1) * sizeof(uint4), false))) return rc;
    if ((rc = entries.reserve(c, 2 * est_nodes * 8, false)) || (rc = leaf_refs.reserve(c, est_refs * 4, false)))
        return rc;
    HIPC(c, hipMemsetAsync(entries.p, 0, 2 * 8, st));
    hipLaunchKernelGGL(prb::k_sah_root, dim3(nblocks(n, 256)), dim3(256), 0, st, box, n, nodes[0].as<LNode>(),
                       refs[0].as<uint32_t>(), rnode[0].as<uint32_t>());
    uint32_t n_nodes = 1, n_refs = n, pair_base = 1, leaf_base = 0;
    const uint64_t ref_cap = 16ull * n + 4096;  // duplication beyond this is refused
    for (uint32_t level = 0;; level++) {
        if ((rc = hist.reserve(c, (size_t)n_nodes * prb::SAH_HIST * 4, false)) ||
            (rc = axis.reserve(c, (size_t)n_nodes * 4, false)) || (rc = pos.reserve(c, (size_t)n_nodes * 4, false)) ||
            (rc = f.reserve(c, ((size_t)n_refs + 1) * sizeof(uint2), false)) ||
            (rc = s.reserve(c, ((size_t)n_refs + 1) * sizeof(uint2), false)) ||
            (rc = v.reserve(c, ((size_t)n_nodes + 1) * sizeof(uint4), false)) ||
            (rc = sn.reserve(c, ((size_t)n_nodes + 1) * sizeof(uint4), false)))
            return rc;
        const LNode *N = nodes[cur].as<LNode>();
        const uint32_t *R = refs[cur].as<uint32_t>(), *RN = rnode[cur].as<uint32_t>();
        // small nodes are decided here; larger ones get a cleared histogram
        hipLaunchKernelGGL(prb::k_sah_small, dim3(nblocks(n_nodes, 4)), dim3(256), 0, st, n_nodes, N, R, tlo, thi, P,
                           level, axis.as<int>(), pos.as<float>(), hist.as<uint32_t>());
        if (n_refs)
            hipLaunchKernelGGL(prb::k_sah_bins, dim3((n_refs + prb::SAH_CHUNK - 1) / prb::SAH_CHUNK), dim3(256), 0,
                               st, n_refs, R, RN, tlo, thi, N, hist.as<uint32_t>());
        hipLaunchKernelGGL(prb::k_sah_split, dim3(nblocks(n_nodes, 256)), dim3(256), 0, st, n_nodes, N,
                           (const uint32_t *)hist.as<uint32_t>(), P, level, axis.as<int>(), pos.as<float>());
        hipLaunchKernelGGL(prb::k_sah_classify, dim3(nblocks((uint64_t)n_refs + 1, 256)), dim3(256), 0, st, n_refs, R,
                           RN, tlo, thi, N, (const int *)axis.as<int>(), (const float *)pos.as<float>(),
                           f.as<uint2>());
        HIPC(c, hipGetLastError());
        if ((rc = scan_sum(c, st, tmp, (const uint2 *)f.as<uint2>(), s.as<uint2>(), (size_t)n_refs + 1,
                           make_uint2(0u, 0u), AddU2{})))
            return rc;
        hipLaunchKernelGGL(prb::k_sah_counts, dim3(nblocks((uint64_t)n_nodes + 1, 256)), dim3(256), 0, st, n_nodes, N,
                           (const int *)axis.as<int>(), (const uint2 *)s.as<uint2>(), v.as<uint4>());
        HIPC(c, hipGetLastError());
        if ((rc = scan_sum(c, st, tmp, (const uint4 *)v.as<uint4>(), sn.as<uint4>(), (size_t)n_nodes + 1,
                           make_uint4(0u, 0u, 0u, 0u), AddU4{})))
            return rc;
        uint4 tot;  // the totals of the level: the scan's entry at n_nodes
        HIPC(c, hipMemcpyAsync(&tot, sn.as<uint4>() + n_nodes, sizeof(uint4), hipMemcpyDeviceToHost, st));
        HIPC(c, hipStreamSynchronize(st));
        const uint32_t n_split = tot.x, next_refs = tot.y, level_leaf = tot.z;
        if ((uint64_t)next_refs + leaf_base + level_leaf > ref_cap)
            return fail(c, PRAY_ERR_UNSUPPORTED, "SAH build: triangle references exceed 16 per triangle");
        const int nxt = cur ^ 1;
        if ((rc = entries.reserve(c, 2 * ((size_t)pair_base + n_split) * 8, true)) ||
            (rc = leaf_refs.reserve(c, ((size_t)leaf_base + level_leaf) * 4, true)) ||
            (rc = nodes[nxt].reserve(c, std::max<size_t>(2 * (size_t)n_split, 1) * sizeof(LNode), false)) ||
            (rc = refs[nxt].reserve(c, std::max<size_t>(next_refs, 1) * 4, false)) ||
            (rc = rnode[nxt].reserve(c, std::max<size_t>(next_refs, 1) * 4, false)))
            return rc;
        prb::Emit E{n_nodes, N, (const int *)axis.as<int>(), (const float *)pos.as<float>(),
                    (const uint2 *)s.as<uint2>(), (const uint4 *)sn.as<uint4>(), pair_base, leaf_base,
                    entries.as<uint2>(), nodes[nxt].as<LNode>()};
        hipLaunchKernelGGL(prb::k_sah_emit, dim3(nblocks(n_nodes, 256)), dim3(256), 0, st, E);
        prb::Scatter Sc{n_refs, R, RN, N, (const int *)axis.as<int>(), (const uint2 *)f.as<uint2>(),
                        (const uint2 *)s.as<uint2>(), (const uint4 *)sn.as<uint4>(), leaf_base,
                        refs[nxt].as<uint32_t>(), rnode[nxt].as<uint32_t>(), leaf_refs.as<uint32_t>()};
        if (n_refs) hipLaunchKernelGGL(prb::k_sah_scatter, dim3(nblocks(n_refs, 256)), dim3(256), 0, st, Sc);
        HIPC(c, hipGetLastError());
        leaf_base += level_leaf;
        pair_base += n_split;
        if (n_split == 0) {
            out.depth = level;
            break;
        }
        cur = nxt;
        n_nodes = 2 * n_split;
        n_refs = next_refs;
    }
    HIPC(c, hipStreamSynchronize(st));
    out.n_entries = 2 * pair_base;
    out.n_refs = leaf_base;
    return PRAY_OK;
}

}  // namespace
}  // namespace prc

using namespace prc;

extern "C" {

int pray_build_scene(pray_ctx *c, const float *tris, const float *color, const uint8_t *emit, uint32_t n, int method,
                     uint32_t leaf_size, uint32_t max_depth, float traversal_cost, float intersection_cost,
                     int *scene_id, float *build_ms) {
    if (!c || !scene_id) return fail(c, PRAY_ERR_INVALID, "null argument");
    *scene_id = -1;
    if (n && !tris) return fail(c, PRAY_ERR_INVALID, "null triangles");
    if (method != PRAY_BUILD_MORTON && method != PRAY_BUILD_SAH)
        return fail(c, PRAY_ERR_INVALID, "unknown build method %d", method);
    if (leaf_size < 1 || leaf_size > 64) return fail(c, PRAY_ERR_INVALID, "leaf_size must be in 1..64");
    if (max_depth == 0) max_depth = 40;
    if (max_depth > MAX_TREE_DEPTH) return fail(c, PRAY_ERR_INVALID, "max_depth must be at most %u", MAX_TREE_DEPTH);
    if (n > (1u << 28)) return fail(c, PRAY_ERR_INVALID, "at most 2^28 triangles");
    const prb::SahParams P{leaf_size, max_depth, traversal_cost > 0.f ? traversal_cost : 2.0f,
                           intersection_cost > 0.f ? intersection_cost : 1.0f};
    HIPC(c, hipSetDevice(c->device));
    hipStream_t st = c->stream;
    auto sc = std::make_unique<SceneRec>();
    // frees the scene's device arrays on every failure return
    struct Release {
        std::unique_ptr<SceneRec> &s;
        ~Release() {
            if (s) free_scene_rec(*s);
        }
    } release{sc};
    prt::DevScene &d = sc->d;
    d.n_tris = n;
    if (n == 0) {
        *scene_id = c->next_id++;
        c->scenes[*scene_id] = std::move(sc);
        if (build_ms) *build_ms = 0.f;
        return PRAY_OK;
    }
    for (int k = 0; k < 2; k++)
        if (!c->ievents[k]) HIPC(c, hipEventCreate(&c->ievents[k]));
    int rc = PRAY_OK;
    Pool &S = build_scratch(c).prep;
    S.next = 0;
    float *abc = nullptr, *tlo = nullptr, *thi = nullptr, *partial = nullptr, *box = nullptr;
    if ((rc = S.get(c, &abc, 9 * (size_t)n)) || (rc = S.get(c, &tlo, 3 * (size_t)n)) ||
        (rc = S.get(c, &thi, 3 * (size_t)n)) || (rc = S.get(c, &partial, 6 * (size_t)prb::BOX_PARTIALS)) ||
        (rc = S.get(c, &box, 6)))
        return rc;
    // Scene arrays (owned by the scene record).
    void *p = nullptr;
    auto own = [&](size_t bytes, void **dst) -> int { return device_alloc(c, *sc, dst, bytes); };
    float *a = nullptr, *ab = nullptr, *ac = nullptr, *normal = nullptr, *col = nullptr;
    uint8_t *em = nullptr;
    if ((rc = own(12 * (size_t)n, &p))) return rc;
    a = (float *)p;
    if ((rc = own(12 * (size_t)n, &p))) return rc;
    ab = (float *)p;
    if ((rc = own(12 * (size_t)n, &p))) return rc;
    ac = (float *)p;
    if (color) {
        if ((rc = own(12 * (size_t)n, &p))) return rc;
        normal = (float *)p;
        if ((rc = own(12 * (size_t)n, &p))) return rc;
        col = (float *)p;
        if ((rc = own((size_t)n, &p))) return rc;
        em = (uint8_t *)p;
    }
    HIPC(c, hipMemcpyAsync(abc, tris, 36 * (size_t)n, hipMemcpyHostToDevice, st));
    if (color) {
        HIPC(c, hipMemcpyAsync(col, color, 12 * (size_t)n, hipMemcpyHostToDevice, st));
        if (emit)
            HIPC(c, hipMemcpyAsync(em, emit, (size_t)n, hipMemcpyHostToDevice, st));
        else
            HIPC(c, hipMemsetAsync(em, 0, (size_t)n, st));
    }

    // ---- timed build: from the uploaded triangles to the device layout ----
    HIPC(c, hipEventRecord(c->ievents[0], st));
    const unsigned g = nblocks(n, 256);
    prb::Tri T{abc, a, ab, ac, normal, tlo, thi};
    hipLaunchKernelGGL(prb::k_prep, dim3(g), dim3(256), 0, st, n, T);
    const unsigned nb = std::min<unsigned>(prb::BOX_PARTIALS, nblocks(n, prb::BOX_BLOCK));
    hipLaunchKernelGGL(prb::k_box_partial, dim3(nb), dim3(prb::BOX_BLOCK), 0, st, n, (const float *)tlo,
                       (const float *)thi, partial);
    hipLaunchKernelGGL(prb::k_box_final, dim3(1), dim3(prb::BOX_BLOCK), 0, st, nb, (const float *)partial, box);
    HIPC(c, hipGetLastError());
    Grow &entries = build_scratch(c).entries, &refs = build_scratch(c).leaf_refs;
    Built B;
    rc = method == PRAY_BUILD_MORTON ? build_morton(c, st, n, leaf_size, tlo, thi, box, entries, refs, B)
                                     : build_sah(c, st, n, P, tlo, thi, box, entries, refs, B);
    if (rc) return rc;
    if (B.depth > MAX_TREE_DEPTH)
        return fail(c, PRAY_ERR_UNSUPPORTED, "tree depth %u exceeds %u", B.depth, MAX_TREE_DEPTH);
    // Final arrays: nodes padded to whole 128-byte blocks, references and
    // records padded by two entries for the leaf look-ahead.
    const size_t n_pad = ((size_t)B.n_entries + 15) & ~(size_t)15;
    if ((rc = own(n_pad * 8, &p))) return rc;
    uint2 *nodes = (uint2 *)p;
    if ((rc = own(((size_t)B.n_refs + 2) * 4, &p))) return rc;
    uint32_t *fr = (uint32_t *)p;
    if ((rc = own(((size_t)B.n_refs + 2) * 48, &p))) return rc;
    float4 *recs = (float4 *)p;
    HIPC(c, hipMemsetAsync(nodes, 0, n_pad * 8, st));
    HIPC(c, hipMemsetAsync(fr, 0, ((size_t)B.n_refs + 2) * 4, st));
    HIPC(c, hipMemsetAsync(recs, 0, ((size_t)B.n_refs + 2) * 48, st));
    HIPC(c, hipMemcpyAsync(nodes, entries.p, (size_t)B.n_entries * 8, hipMemcpyDeviceToDevice, st));
    HIPC(c, hipMemcpyAsync(fr, refs.p, (size_t)B.n_refs * 4, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(prb::k_pack_leaf_records, dim3(nblocks(B.n_refs, 256)), dim3(256), 0, st, B.n_refs,
                       (const uint32_t *)fr, (const float *)a, (const float *)ab, (const float *)ac, recs);
    HIPC(c, hipGetLastError());
    d.nodes = nodes;
    d.refs = fr;
    d.a = a, d.ab = ab, d.ac = ac;
    d.normal = normal, d.color = col, d.emit = em;
    d.recs = recs;
    d.n_nodes = B.n_entries;
    d.n_refs = B.n_refs;
    d.max_depth = B.depth;
    // The SAH tree's pairs have equal clips: its two-level node records are
    // part of its layout (the build waits for their check).
    if (method == PRAY_BUILD_SAH && (rc = build_quads(c, *sc, st))) return rc;
    HIPC(c, hipEventRecord(c->ievents[1], st));
    // ---- end of the timed build ----

    float box_h[6];
    HIPC(c, hipMemcpyAsync(box_h, box, 24, hipMemcpyDeviceToHost, st));
    HIPC(c, hipStreamSynchronize(st));
    if (build_ms) HIPC(c, hipEventElapsedTime(build_ms, c->ievents[0], c->ievents[1]));
    std::memcpy(d.box, box_h, sizeof d.box);
    *scene_id = c->next_id++;
    c->scenes[*scene_id] = std::move(sc);
    return PRAY_OK;
}

int pray_scene_layout(pray_ctx *c, int scene_id, uint64_t *sizes, uint32_t *nodes, uint32_t *refs, float *box) {
    if (!c) return PRAY_ERR_INVALID;
    SceneRec *sc = find_scene(c, scene_id);
    if (!sc) return fail(c, PRAY_ERR_NOSCENE, "unknown scene %d", scene_id);
    const prt::DevScene &d = sc->d;
    if (sizes) sizes[0] = d.n_nodes, sizes[1] = d.n_refs, sizes[2] = d.n_tris, sizes[3] = d.max_depth;
    HIPC(c, hipSetDevice(c->device));
    if (nodes && d.n_nodes)
        HIPC(c, hipMemcpyAsync(nodes, d.nodes, (size_t)d.n_nodes * 8, hipMemcpyDeviceToHost, c->stream));
    if (refs && d.n_refs)
        HIPC(c, hipMemcpyAsync(refs, d.refs, (size_t)d.n_refs * 4, hipMemcpyDeviceToHost, c->stream));
    if (box) std::memcpy(box, d.box, sizeof d.box);
    HIPC(c, hipStreamSynchronize(c->stream));
    return PRAY_OK;
}

}  // extern "C"
