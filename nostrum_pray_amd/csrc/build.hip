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
// k_radix_tree, k_bottom_up (boxes, counts, collapsed heights), k_write_nodes
// and k_records.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <memory>
#include <rocprim/device/device_radix_sort.hpp>

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
    uint32_t *parent;         // of inner nodes (n-1 entries; the root's is unused)
    uint32_t *leaf_parent;    // of sorted positions
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
    if (lc & LEAF_BIT) R.leaf_parent[g] = (uint32_t)i; else R.parent[g] = (uint32_t)i;
    if (rc & LEAF_BIT) R.leaf_parent[g + 1] = (uint32_t)i; else R.parent[g + 1] = (uint32_t)i;
}

struct Up {
    uint32_t n, leaf_size;
    const uint32_t *vals;       // sorted position -> triangle
    const float *tlo, *thi;     // triangle boxes
    const uint32_t *left, *right, *parent, *leaf_parent;
    uint32_t *arrivals;         // n-1 counters, zero at launch
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

// One thread per sorted position walks up; the second thread to reach an
// inner node combines its children and goes on.
static __global__ void k_bottom_up(Up U) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= U.n || U.n < 2) return;
    uint32_t node = U.leaf_parent[k];
    while (true) {
        __threadfence();
        if (atomicAdd(&U.arrivals[node], 1u) == 0u) return;  // the sibling is not done yet
        __threadfence();
        float llo[3], lhi[3], rlo[3], rhi[3];
        uint32_t lc, lh, rc, rh;
        child_info(U, U.left[node], llo, lhi, lc, lh);
        child_info(U, U.right[node], rlo, rhi, rc, rh);
        for (int q = 0; q < 3; q++) {
            U.nlo[3 * (size_t)node + q] = fminf(llo[q], rlo[q]);
            U.nhi[3 * (size_t)node + q] = fmaxf(lhi[q], rhi[q]);
        }
        const uint32_t cnt = lc + rc;
        U.count[node] = cnt;
        U.height[node] = cnt <= U.leaf_size ? 0u : 1u + max(lh, rh);
        if (node == 0) return;
        node = U.parent[node];
    }
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

// Leaf references and packed records in sorted order.
static __global__ void k_records(uint32_t n, const uint32_t *vals, const uint8_t *last_flag, const float *a,
                                 const float *ab, const float *ac, uint32_t *refs, float4 *recs) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t t = vals[k];
    const uint32_t ref = t | (last_flag[k] ? REF_LAST : 0u);
    refs[k] = ref;
    const size_t o = 3 * (size_t)t;
    recs[3 * (size_t)k] = make_float4(a[o], a[o + 1], a[o + 2], __uint_as_float(ref));
    recs[3 * (size_t)k + 1] = make_float4(ab[o], ab[o + 1], ab[o + 2], 0.f);
    recs[3 * (size_t)k + 2] = make_float4(ac[o], ac[o + 1], ac[o + 2], 0.f);
}

}  // namespace prb

namespace prc {
namespace {

// Scratch of one build, freed on every exit.
struct Scratch {
    std::vector<void *> ptrs;
    ~Scratch() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <class T> int get(pray_ctx *c, T **p, size_t count) {
        void *q = nullptr;
        HIPC(c, hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
        ptrs.push_back(q);
        *p = static_cast<T *>(q);
        return PRAY_OK;
    }
};

}  // namespace
}  // namespace prc

using namespace prc;

extern "C" {

int pray_build_scene(pray_ctx *c, const float *tris, const float *color, const uint8_t *emit, uint32_t n,
                     uint32_t leaf_size, int *scene_id, float *build_ms) {
    if (!c || !scene_id) return fail(c, PRAY_ERR_INVALID, "null argument");
    *scene_id = -1;
    if (n && !tris) return fail(c, PRAY_ERR_INVALID, "null triangles");
    if (leaf_size < 1 || leaf_size > 64) return fail(c, PRAY_ERR_INVALID, "leaf_size must be in 1..64");
    if (n > (1u << 28)) return fail(c, PRAY_ERR_INVALID, "at most 2^28 triangles");
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
    Scratch S;
    float *abc = nullptr, *tlo = nullptr, *thi = nullptr, *partial = nullptr, *box = nullptr, *nlo = nullptr,
          *nhi = nullptr;
    uint32_t *keys = nullptr, *vals = nullptr, *keys2 = nullptr, *vals2 = nullptr, *left = nullptr,
             *right = nullptr, *parent = nullptr, *leaf_parent = nullptr, *first = nullptr, *last = nullptr,
             *arrivals = nullptr, *count = nullptr, *height = nullptr;
    uint8_t *last_flag = nullptr;
    const size_t inner = n > 1 ? n - 1 : 1;
    if ((rc = S.get(c, &abc, 9 * (size_t)n)) || (rc = S.get(c, &tlo, 3 * (size_t)n)) ||
        (rc = S.get(c, &thi, 3 * (size_t)n)) || (rc = S.get(c, &partial, 6 * (size_t)prb::BOX_PARTIALS)) ||
        (rc = S.get(c, &box, 6)) || (rc = S.get(c, &keys, n)) || (rc = S.get(c, &vals, n)) ||
        (rc = S.get(c, &keys2, n)) || (rc = S.get(c, &vals2, n)) || (rc = S.get(c, &left, inner)) ||
        (rc = S.get(c, &right, inner)) || (rc = S.get(c, &parent, inner)) || (rc = S.get(c, &leaf_parent, n)) ||
        (rc = S.get(c, &first, inner)) || (rc = S.get(c, &last, inner)) || (rc = S.get(c, &arrivals, inner)) ||
        (rc = S.get(c, &count, inner)) || (rc = S.get(c, &height, inner)) || (rc = S.get(c, &nlo, 3 * inner)) ||
        (rc = S.get(c, &nhi, 3 * inner)) || (rc = S.get(c, &last_flag, n)))
        return rc;
    // Scene arrays (owned by the scene record).
    void *p = nullptr;
    const size_t n_nodes = 2 * (size_t)n, n_nodes_pad = (n_nodes + 15) & ~(size_t)15;
    auto own = [&](size_t bytes, void **dst) -> int { return device_alloc(c, *sc, dst, bytes); };
    float *a = nullptr, *ab = nullptr, *ac = nullptr, *normal = nullptr, *col = nullptr;
    uint8_t *em = nullptr;
    uint2 *nodes = nullptr;
    uint32_t *refs = nullptr;
    float4 *recs = nullptr;
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
    if ((rc = own(n_nodes_pad * 8, &p))) return rc;
    nodes = (uint2 *)p;
    if ((rc = own(((size_t)n + 2) * 4, &p))) return rc;
    refs = (uint32_t *)p;
    if ((rc = own(((size_t)n + 2) * 48, &p))) return rc;
    recs = (float4 *)p;

    HIPC(c, hipMemcpyAsync(abc, tris, 36 * (size_t)n, hipMemcpyHostToDevice, st));
    if (color) {
        HIPC(c, hipMemcpyAsync(col, color, 12 * (size_t)n, hipMemcpyHostToDevice, st));
        if (emit)
            HIPC(c, hipMemcpyAsync(em, emit, (size_t)n, hipMemcpyHostToDevice, st));
        else
            HIPC(c, hipMemsetAsync(em, 0, (size_t)n, st));
    }
    HIPC(c, hipMemsetAsync(nodes, 0, n_nodes_pad * 8, st));
    HIPC(c, hipMemsetAsync(refs, 0, ((size_t)n + 2) * 4, st));
    HIPC(c, hipMemsetAsync(recs, 0, ((size_t)n + 2) * 48, st));
    HIPC(c, hipMemsetAsync(arrivals, 0, inner * 4, st));
    HIPC(c, hipMemsetAsync(last_flag, 0, (size_t)n, st));
    // rocPRIM's scratch size for this count
    size_t sort_bytes = 0;
    HIPC(c, rocprim::radix_sort_pairs(nullptr, sort_bytes, keys, keys2, vals, vals2, n, 0u, prb::MORTON_BITS, st));
    uint8_t *sort_tmp = nullptr;
    if ((rc = S.get(c, &sort_tmp, sort_bytes))) return rc;

    // ---- timed build ----
    HIPC(c, hipEventRecord(c->ievents[0], st));
    const unsigned g = nblocks(n, 256);
    prb::Tri T{abc, a, ab, ac, normal, tlo, thi};
    hipLaunchKernelGGL(prb::k_prep, dim3(g), dim3(256), 0, st, n, T);
    const unsigned nb = std::min<unsigned>(prb::BOX_PARTIALS, nblocks(n, prb::BOX_BLOCK));
    hipLaunchKernelGGL(prb::k_box_partial, dim3(nb), dim3(prb::BOX_BLOCK), 0, st, n, (const float *)tlo,
                       (const float *)thi, partial);
    hipLaunchKernelGGL(prb::k_box_final, dim3(1), dim3(prb::BOX_BLOCK), 0, st, nb, (const float *)partial, box);
    hipLaunchKernelGGL(prb::k_morton, dim3(g), dim3(256), 0, st, n, (const float *)tlo, (const float *)thi,
                       (const float *)box, keys, vals);
    HIPC(c, hipGetLastError());
    HIPC(c, rocprim::radix_sort_pairs(sort_tmp, sort_bytes, keys, keys2, vals, vals2, n, 0u, prb::MORTON_BITS, st));
    prb::Radix R{keys2, n, left, right, parent, leaf_parent, first, last};
    prb::Up U{n, leaf_size, vals2, tlo, thi, left, right, parent, leaf_parent, arrivals, nlo, nhi, count, height};
    prb::Writer W{n, leaf_size, keys2, left, right, first, last, count, nlo, nhi, nodes, last_flag};
    if (n > 1) {
        hipLaunchKernelGGL(prb::k_radix_tree, dim3(nblocks(n - 1, 256)), dim3(256), 0, st, R);
        hipLaunchKernelGGL(prb::k_bottom_up, dim3(g), dim3(256), 0, st, U);
    }
    hipLaunchKernelGGL(prb::k_write_nodes, dim3(nblocks(inner, 256)), dim3(256), 0, st, W, U);
    hipLaunchKernelGGL(prb::k_records, dim3(g), dim3(256), 0, st, n, (const uint32_t *)vals2,
                       (const uint8_t *)last_flag, (const float *)a, (const float *)ab, (const float *)ac, refs, recs);
    HIPC(c, hipGetLastError());
    HIPC(c, hipEventRecord(c->ievents[1], st));
    // ---- end of the timed build ----

    uint32_t depth = 0;
    float box_h[6];
    if (n > 1) HIPC(c, hipMemcpyAsync(&depth, height, 4, hipMemcpyDeviceToHost, st));
    HIPC(c, hipMemcpyAsync(box_h, box, 24, hipMemcpyDeviceToHost, st));
    HIPC(c, hipStreamSynchronize(st));
    if (build_ms) HIPC(c, hipEventElapsedTime(build_ms, c->ievents[0], c->ievents[1]));
    if (depth > MAX_TREE_DEPTH)
        return fail(c, PRAY_ERR_UNSUPPORTED, "tree depth %u exceeds %u", depth, MAX_TREE_DEPTH);
    d.nodes = nodes;
    d.refs = refs;
    d.a = a, d.ab = ab, d.ac = ac;
    d.normal = normal, d.color = col, d.emit = em;
    d.recs = recs;
    std::memcpy(d.box, box_h, sizeof d.box);
    d.n_nodes = (uint32_t)n_nodes;
    d.n_refs = n;
    d.max_depth = depth;
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
