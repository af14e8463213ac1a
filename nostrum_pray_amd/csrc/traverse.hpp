// Closest-hit / any-hit traversal kernel for gfx950.
//
// One lane = one ray at a time; waves are persistent and refill lanes whose
// ray finished from the wave's pool of ray indices, so long rays never hold a
// wave of finished lanes.  The pool takes rays in chunks of 64 from NCTR
// counters 256 bytes apart: counter j hands out chunks j, j + NCTR, ...; a wave
// starts at counter (global wave index % NCTR) and moves on to the next when
// one runs dry.  Thousands of resident waves taking every refill from one
// counter serialise on its address (an atomic per refill); here each address
// sees 1/NCTR of one atomic per 64 rays, and no wave is assumed to be resident
// from the start.
//
// Every loop iteration each lane does an inner-node step (one aligned 16-byte
// child-pair load), a leaf step (ILP triangle records loaded together; the
// record array is padded so the look-ahead stays in bounds), or both when the
// node step lands in a leaf.  Pending far children go to a per-lane stack: K
// entries in an LDS ring, the oldest entries spilled to a per-lane global area
// when the ring is full.
//
// Interval arithmetic: planes are intersected front to back with the
// direction's sign choosing near/far; every plane distance is widened by a
// relative 1e-6 (conservative: a cell is never skipped by rounding; infinite
// distances stay infinite), and a pending cell is pruned only when its widened
// entry exceeds the best t (strictly), so equal-t ties are always examined.  The closest hit is the
// lexicographic minimum of (t, triangle index) over t > 0, which equals the
// reference brute force (bf_search.cpp:26-40) on every ray.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cfloat>

#include "device_math.hpp"

namespace prt {

constexpr int WAVE = 64;
constexpr uint32_t LEAF_TAG = 3u;
constexpr uint32_t LEAF_EMPTY = 0x3FFFFFFFu;
constexpr uint32_t REF_LAST = 0x80000000u;
constexpr uint32_t NO_HIT = 0xFFFFFFFFu;
constexpr float WIDEN = 1e-6f;
constexpr int MAX_WG_WAVES = 4;      // waves per workgroup (each an independent persistent wave)
constexpr uint32_t NCTR = 64;        // ray-chunk counters
constexpr uint32_t CTR_STRIDE = 64;  // uint32 words between counters (256 bytes)

struct DevScene {
    const uint2 *nodes = nullptr;
    const uint32_t *refs = nullptr;   // n_refs + 2 (padded)
    const float *a = nullptr, *ab = nullptr, *ac = nullptr;
    const float4 *recs = nullptr;     // 3 * (n_refs + 2) (padded)
    const uint4 *quads = nullptr;     // 2 per children pair: two-level node records (SAH trees), or null
    const uint4 *boxes = nullptr;     // 4 per pair of the boxed layout (median trees), or null
    const float4 *brecs = nullptr;    // packed records of the boxed layout's leaf runs (padded)
    const uint4 *boxes4 = nullptr;    // 8 per record of the 4-wide boxed layout, or null
    const float *normal = nullptr, *color = nullptr;
    const uint8_t *emit = nullptr;
    float box[6] = {0, 0, 0, 0, 0, 0};
    uint32_t n_tris = 0, n_nodes = 0, n_refs = 0, max_depth = 0;
    uint32_t box_depth = 0;           // depth of the boxed layout
    uint32_t box4_depth = 0, box4_root = 0;  // stack entries and root word of the 4-wide layout
};

struct TraceArgs {
    const float *rays;        // n x 6
    uint32_t n;               // rays (or launch bound when n_dev is set)
    const uint32_t *n_dev;    // optional device-side ray count
    const float *limit2;      // any-hit squared limits
    float eps;                // determinant cull
    int inclusive;            // any-hit: t*t <= limit2 instead of <
    int32_t *out_idx;
    float *out_t;             // optional
    uint32_t *counts;         // optional: 2 per ray (nodes, refs)
    uint32_t *work;           // NCTR chunk counters, CTR_STRIDE apart (0 at launch)
    uint4 *spill;             // per lane slot: spill_depth entries
    uint32_t spill_depth;
    uint32_t fetch_min;       // refill when at least this many lanes are idle
    uint32_t leaf_batch;      // 0: node and leaf steps in every iteration; T > 0: an iteration takes
                              // leaf steps when at least T lanes wait at a leaf (or no lane needs a
                              // node step), node steps otherwise
    const DevScene *table;    // multi-scene launches: scene of ray i is
    const uint32_t *ray_scene;  // table[ray_scene[i]]
};

enum : uint32_t { ST_IDLE = 0, ST_NODE = 1, ST_LEAF = 2, ST_DEAD = 3 };

// Widening by a relative WIDEN, multiplicatively so that +-infinity (a zero
// direction component) stays infinite: x +- |x| * WIDEN would turn +inf into
// inf - inf = NaN, which fminf/fmaxf ignore, and such a ray would lose all
// pruning along that axis and walk most of the tree.
__device__ __forceinline__ float whi(float x) { return x * (x > 0.f ? 1.f + WIDEN : 1.f - WIDEN); }
__device__ __forceinline__ float wlo(float x) { return x * (x > 0.f ? 1.f - WIDEN : 1.f + WIDEN); }
__device__ __forceinline__ bool empty_leaf(uint32_t w) { return w == ((LEAF_EMPTY << 2) | LEAF_TAG); }

__device__ __forceinline__ void slab(float bmin, float bmax, float o, float inv, float &lo, float &hi) {
    float t0 = (bmin - o) * inv, t1 = (bmax - o) * inv;
    if (inv < 0.f) {
        const float s = t0;
        t0 = t1;
        t1 = s;
    }
    lo = fmaxf(lo, wlo(t0));
    hi = fminf(hi, whi(t1));
}

// The values must be in registers at this point: placed after a leaf step's
// record loads, it keeps the compiler from sinking a load below the branch
// that first uses it (a triangle's first vertex is only read after its
// determinant test, which would cost the step a second memory round trip),
// so all loads of the step are in flight together.
__device__ __forceinline__ void hold(const float4 &v) { asm volatile("" ::"v"(v.x), "v"(v.y), "v"(v.z)); }
__device__ __forceinline__ void hold(const uint4 &v) { asm volatile("" ::"v"(v.x), "v"(v.y), "v"(v.z), "v"(v.w)); }

template <int LAYOUT>
__device__ __forceinline__ void load_tri(const DevScene &s, const float4 *recs, uint32_t k, uint32_t &ref, float4 &A,
                                         float4 &B, float4 &Cc) {
    if (LAYOUT == 1) {
        A = recs[3 * (size_t)k];
        B = recs[3 * (size_t)k + 1];
        Cc = recs[3 * (size_t)k + 2];
        ref = __float_as_uint(A.w);
    } else {
        ref = s.refs[k];
        const size_t t = 3 * (size_t)(ref & ~REF_LAST);
        A = make_float4(s.a[t], s.a[t + 1], s.a[t + 2], 0.f);
        B = make_float4(s.ab[t], s.ab[t + 1], s.ab[t + 2], 0.f);
        Cc = make_float4(s.ac[t], s.ac[t + 1], s.ac[t + 2], 0.f);
    }
}

// MULTI: rays of several scenes in one launch (packed layout only); each lane
// takes its ray's node and record arrays and root cell from the scene table
// when it picks the ray up.
//
// A workgroup holds 1..MAX_WG_WAVES waves that never synchronise with each
// other; each has its own LDS ring in the dynamic shared memory the launch
// sizes (ring_lds_bytes per wave).
// BATCH: leaf batching (TraceArgs::leaf_batch > 0) compiled in; without it
// every iteration takes node and leaf steps and the loop carries no batching
// code.
// BOX == 1: node steps over the boxed layout (DevScene::boxes, with the records
// of its leaf runs in DevScene::brecs) instead of the 8-byte nodes' clip
// planes.  Per
// pair p, boxes[4p + a] for axis a = 0, 1, 2 holds {child 0's lower bound,
// child 1's lower bound, child 0's upper bound, child 1's upper bound} along
// that axis (float bits) and boxes[4p + 3] the children's words (an inner
// child's word names its pair as w >> 3; pair 0's child 1 is the root).  A
// step intersects the ray's interval, capped at the best t, with both
// children's boxes, enters the child with the smaller entry distance (child
// 0 on a tie) and stacks the other.  For trees whose children overlap
// (median trees: the children's boxes are those of their triangles) this
// prunes on all three axes where the clip planes bound one.
// BOX == 2: the same over the 4-wide form (DevScene::boxes4, root word
// DevScene::box4_root): record q is boxes4[8q .. 8q + 7] = {lower x of slots
// 0..3}, {upper x}, {lower y}, {upper y}, {lower z}, {upper z}, {the slots'
// words}, unused.  A step intersects the ray's interval with the four
// slots' boxes, orders the live ones by entry distance (lower slot first on
// a tie), enters the nearest and stacks the others farthest first, so that
// pops return them nearest first.
template <bool ANY, int LAYOUT, int K, int ILP, bool MULTI = false, bool BATCH = false, int BOX = 0>
static __global__ __launch_bounds__(MAX_WG_WAVES * WAVE) void k_trace(DevScene s, TraceArgs A) {
    static_assert(K >= 2 && K <= 16, "LDS ring of 2..16 entries");
    static_assert(!MULTI || LAYOUT == 1, "multi-scene launches use the packed layout");
    static_assert(BOX == 0 || (LAYOUT == 1 && !MULTI), "boxed node steps trace the packed layout of one scene");
    extern __shared__ uint32_t ring_lds[];
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t wave_in_block = threadIdx.x / WAVE;
    const uint32_t gwave = blockIdx.x * (blockDim.x / WAVE) + wave_in_block;  // global wave index
    uint32_t *st_w = ring_lds + wave_in_block * (3 * K * WAVE);
    float *st_lo = reinterpret_cast<float *>(st_w + K * WAVE);
    float *st_hi = reinterpret_cast<float *>(st_w + 2 * K * WAVE);
    const uint64_t below = (1ull << lane) - 1ull;
    uint4 *spill = A.spill + ((size_t)gwave * WAVE + lane) * A.spill_depth;
    const uint32_t nn = A.n_dev ? *A.n_dev : A.n;
    const uint4 *pairs = reinterpret_cast<const uint4 *>(s.nodes);
    const float4 *recs = BOX ? s.brecs : s.recs;
    // The root's word (node 1, the second word pair of the first 16 bytes;
    // child 1 of the boxed layout's pair 0; the 4-wide layout's root word),
    // read once per wave, not once per ray (single-scene launches).
    const uint32_t root_w =
        MULTI ? 0u
              : (BOX == 2 ? s.box4_root
                          : (uint32_t)__builtin_amdgcn_readfirstlane((int)(BOX ? s.boxes[3].y : pairs[0].w)));

    uint32_t state = ST_IDLE, ray = 0, w = 0, k = 0, best_i = NO_HIT, cn = 0, cr = 0;
    float ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0, ix = 0, iy = 0, iz = 0, lim2 = 0;
    float lo = 0, hi = 0, best_t = INFINITY;
    int sp = 0, base = 0;
    // Wave-uniform pool of ray indices [pool_next, pool_end) and the counter
    // the wave takes its next chunk from.
    const uint32_t nchunks = (nn + WAVE - 1) / WAVE;
    uint32_t pool_next = 0, pool_end = 0, ctr = gwave % NCTR;

    while (true) {
        // ------------------------------------------------------- refill --
        const uint64_t idle = __ballot(state == ST_IDLE);
        const uint64_t dead = __ballot(state == ST_DEAD);
        if (dead == ~0ull) break;
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (nidle && (nidle >= A.fetch_min || (idle | dead) == ~0ull)) {
            // The pool holds fewer rays than idle lanes: the next chunk
            // (one chunk covers any shortfall, nidle <= 64).
            const uint32_t avail = pool_end - pool_next;
            uint32_t nb = 0;
            if (avail < nidle) {
                // One lane walks the counters from the current one until a
                // chunk is left (a relaxed read skips drained counters without
                // an atomic); none left: nb points past the last ray.
                const uint32_t first = (uint32_t)(__ffsll((unsigned long long)idle) - 1);
                uint32_t c = nchunks;
                if (lane == first) {
                    for (uint32_t tries = 0; tries < NCTR; ++tries) {
                        const uint32_t own = ctr < nchunks ? (nchunks - ctr + NCTR - 1) / NCTR : 0u;
                        uint32_t *p = A.work + ctr * CTR_STRIDE;
                        if (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < own) {
                            const uint32_t i = atomicAdd(p, 1u);
                            if (i < own) {
                                c = ctr + i * NCTR;
                                break;
                            }
                        }
                        ctr = ctr + 1 == NCTR ? 0u : ctr + 1;
                    }
                }
                // wave-uniform from here: scalar registers
                c = (uint32_t)__builtin_amdgcn_readlane((int)c, (int)first);
                ctr = (uint32_t)__builtin_amdgcn_readlane((int)ctr, (int)first);
                nb = c * WAVE;
            }
            const uint32_t r = (uint32_t)__popcll(idle & below);
            if (state == ST_IDLE) {
                ray = r < avail ? pool_next + r : nb + (r - avail);
                if (ray >= nn) {
                    state = ST_DEAD;
                } else {
                    const float2 r0 = reinterpret_cast<const float2 *>(A.rays)[3 * (size_t)ray];
                    const float2 r1 = reinterpret_cast<const float2 *>(A.rays)[3 * (size_t)ray + 1];
                    const float2 r2 = reinterpret_cast<const float2 *>(A.rays)[3 * (size_t)ray + 2];
                    ox = r0.x, oy = r0.y, oz = r1.x, dx = r1.y, dy = r2.x, dz = r2.y;
                    ix = 1.0f / dx, iy = 1.0f / dy, iz = 1.0f / dz;
                    lo = 0.f, hi = INFINITY;
                    if (MULTI) {
                        const DevScene &S = A.table[A.ray_scene[ray]];
                        pairs = reinterpret_cast<const uint4 *>(S.nodes);
                        recs = S.recs;
                        slab(S.box[0], S.box[3], ox, ix, lo, hi);
                        slab(S.box[1], S.box[4], oy, iy, lo, hi);
                        slab(S.box[2], S.box[5], oz, iz, lo, hi);
                    } else {
                        slab(s.box[0], s.box[3], ox, ix, lo, hi);
                        slab(s.box[1], s.box[4], oy, iy, lo, hi);
                        slab(s.box[2], s.box[5], oz, iz, lo, hi);
                    }
                    best_t = INFINITY, best_i = NO_HIT, sp = 0, base = 0, cn = 0, cr = 0;
                    bool live = true;
                    if (ANY) {
                        lim2 = A.limit2[ray];
                        if (!(lim2 >= 0.f))
                            live = false;
                        else
                            hi = fminf(hi, whi(sqrtf(lim2)));
                    }
                    if (live && lo <= hi) {
                        cn = 1;
                        w = MULTI ? pairs[0].w : root_w;  // node 1: the root
                        if ((w & 3u) != LEAF_TAG) {
                            state = ST_NODE;
                        } else if (!empty_leaf(w)) {
                            state = ST_LEAF;
                            k = w >> 2;
                        } else {
                            live = false;
                        }
                    } else {
                        live = false;
                    }
                    if (!live) {
                        A.out_idx[ray] = -1;
                        if (A.out_t) A.out_t[ray] = FLT_MAX;
                        if (A.counts) A.counts[2 * (size_t)ray] = cn, A.counts[2 * (size_t)ray + 1] = 0;
                        state = ST_IDLE;
                    }
                }
            }
            if (avail < nidle) {
                pool_next = nb + (nidle - avail);
                pool_end = nb + WAVE;
            } else {
                pool_next += nidle;
            }
            pool_next = (uint32_t)__builtin_amdgcn_readfirstlane((int)pool_next);
            pool_end = (uint32_t)__builtin_amdgcn_readfirstlane((int)pool_end);
        }

        // Which steps this iteration takes (wave-uniform): both, or with
        // leaf batching the leaf steps once enough lanes wait at a leaf.
        bool run_node = true, run_leaf = true;
        if constexpr (BATCH) {
            const uint64_t at_leaf = __ballot(state == ST_LEAF);
            const uint64_t at_node = __ballot(state == ST_NODE);
            run_leaf = at_leaf != 0ull && ((uint32_t)__popcll(at_leaf) >= A.leaf_batch || at_node == 0ull);
            run_node = !run_leaf;
        }
        bool need_pop = false;
        // --------------------------------------- 4-wide boxed node step --
        if constexpr (BOX == 2) {
            if (run_node && state == ST_NODE) {
                const uint4 *bp = s.boxes4 + 8 * (size_t)(w >> 3);
                const uint4 qlx = bp[0], qhx = bp[1], qly = bp[2], qhy = bp[3], qlz = bp[4], qhz = bp[5];
                const uint4 qw = bp[6];
                cn += 4;
                const float cap = ANY ? hi : fminf(hi, best_t);
                float l[4], h[4];
                uint32_t cw[4];
                const uint32_t lxa[4] = {qlx.x, qlx.y, qlx.z, qlx.w}, hxa[4] = {qhx.x, qhx.y, qhx.z, qhx.w};
                const uint32_t lya[4] = {qly.x, qly.y, qly.z, qly.w}, hya[4] = {qhy.x, qhy.y, qhy.z, qhy.w};
                const uint32_t lza[4] = {qlz.x, qlz.y, qlz.z, qlz.w}, hza[4] = {qhz.x, qhz.y, qhz.z, qhz.w};
                cw[0] = qw.x, cw[1] = qw.y, cw[2] = qw.z, cw[3] = qw.w;
                // key: entry distance, +inf for a slot not entered
                float key[4];
                uint32_t sl[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    l[j] = lo, h[j] = cap;
                    slab(__uint_as_float(lxa[j]), __uint_as_float(hxa[j]), ox, ix, l[j], h[j]);
                    slab(__uint_as_float(lya[j]), __uint_as_float(hya[j]), oy, iy, l[j], h[j]);
                    slab(__uint_as_float(lza[j]), __uint_as_float(hza[j]), oz, iz, l[j], h[j]);
                    key[j] = (l[j] <= h[j] && !empty_leaf(cw[j])) ? l[j] : INFINITY;
                    sl[j] = (uint32_t)j;
                }
                // sorting network over (key, slot), ties by slot
                auto cx = [&](int a, int b) {
                    const bool sw = key[b] < key[a] || (key[b] == key[a] && sl[b] < sl[a]);
                    const float ka = key[a], kb = key[b];
                    const uint32_t sa = sl[a], sb = sl[b];
                    key[a] = sw ? kb : ka, key[b] = sw ? ka : kb;
                    sl[a] = sw ? sb : sa, sl[b] = sw ? sa : sb;
                };
                cx(0, 1), cx(2, 3), cx(0, 2), cx(1, 3), cx(1, 2);
                auto pick = [&](uint32_t j, uint32_t &pw, float &pl, float &ph) {
                    pw = j == 0 ? cw[0] : (j == 1 ? cw[1] : (j == 2 ? cw[2] : cw[3]));
                    pl = j == 0 ? l[0] : (j == 1 ? l[1] : (j == 2 ? l[2] : l[3]));
                    ph = j == 0 ? h[0] : (j == 1 ? h[1] : (j == 2 ? h[2] : h[3]));
                };
#pragma unroll
                for (int q = 3; q >= 1; --q) {
                    if (key[q] != INFINITY) {
                        uint32_t pw;
                        float pl, ph;
                        pick(sl[q], pw, pl, ph);
                        if (sp - base == K) {  // ring full: spill the oldest entry
                            const uint32_t at = ((uint32_t)base % K) * WAVE + lane;
                            spill[base] =
                                make_uint4(st_w[at], __float_as_uint(st_lo[at]), __float_as_uint(st_hi[at]), 0u);
                            ++base;
                        }
                        const uint32_t at = ((uint32_t)sp % K) * WAVE + lane;
                        st_w[at] = pw;
                        st_lo[at] = pl;
                        st_hi[at] = ph;
                        ++sp;
                    }
                }
                if (key[0] != INFINITY) {
                    pick(sl[0], w, lo, hi);
                    if ((w & 3u) == LEAF_TAG) {
                        state = ST_LEAF;
                        k = w >> 2;
                    }
                } else {
                    need_pop = true;
                }
            }
        }
        // ---------------------------------------------- boxed node step --
        if constexpr (BOX == 1) {
            if (run_node && state == ST_NODE) {
                const uint4 *bp = s.boxes + 4 * (size_t)(w >> 3);
                const uint4 qx = bp[0], qy = bp[1], qz = bp[2], qw = bp[3];
                cn += 2;
                const float cap = ANY ? hi : fminf(hi, best_t);
                float l0 = lo, h0 = cap, l1 = lo, h1 = cap;
                slab(__uint_as_float(qx.x), __uint_as_float(qx.z), ox, ix, l0, h0);
                slab(__uint_as_float(qx.y), __uint_as_float(qx.w), ox, ix, l1, h1);
                slab(__uint_as_float(qy.x), __uint_as_float(qy.z), oy, iy, l0, h0);
                slab(__uint_as_float(qy.y), __uint_as_float(qy.w), oy, iy, l1, h1);
                slab(__uint_as_float(qz.x), __uint_as_float(qz.z), oz, iz, l0, h0);
                slab(__uint_as_float(qz.y), __uint_as_float(qz.w), oz, iz, l1, h1);
                const bool g0 = l0 <= h0 && !empty_leaf(qw.x);
                const bool g1 = l1 <= h1 && !empty_leaf(qw.y);
                const bool n1 = g1 && (!g0 || l1 < l0);
                if (g0 && g1) {
                    if (sp - base == K) {  // ring full: spill the oldest entry
                        const uint32_t at = ((uint32_t)base % K) * WAVE + lane;
                        spill[base] =
                            make_uint4(st_w[at], __float_as_uint(st_lo[at]), __float_as_uint(st_hi[at]), 0u);
                        ++base;
                    }
                    const uint32_t at = ((uint32_t)sp % K) * WAVE + lane;
                    st_w[at] = n1 ? qw.x : qw.y;
                    st_lo[at] = n1 ? l0 : l1;
                    st_hi[at] = n1 ? h0 : h1;
                    ++sp;
                }
                if (g0 || g1) {
                    w = n1 ? qw.y : qw.x;
                    lo = n1 ? l1 : l0;
                    hi = n1 ? h1 : h0;
                    if ((w & 3u) == LEAF_TAG) {
                        state = ST_LEAF;
                        k = w >> 2;
                    }
                } else {
                    need_pop = true;
                }
            }
        }
        // ---------------------------------------------------- node step --
        if (BOX == 0 && run_node && state == ST_NODE) {
            const uint32_t axis = w & 3u;
            const uint4 pr = pairs[w >> 3];  // pair index = (w >> 2) >> 1
            cn += 2;
            const float oa = axis == 0 ? ox : (axis == 1 ? oy : oz);
            const float ia = axis == 0 ? ix : (axis == 1 ? iy : iz);
            const float tL = (__uint_as_float(pr.x) - oa) * ia;
            const float tR = (__uint_as_float(pr.z) - oa) * ia;
            const bool neg = ia < 0.f;
            const float tn = neg ? tR : tL, tf = neg ? tL : tR;
            const uint32_t wn = neg ? pr.w : pr.y, wf = neg ? pr.y : pr.w;
            const float cap = ANY ? hi : fminf(hi, best_t);
            const float n_hi = fminf(cap, whi(tn));
            const float f_lo = fmaxf(lo, wlo(tf));
            const bool gn = lo <= n_hi && !empty_leaf(wn);
            const bool gf = f_lo <= cap && !empty_leaf(wf);
            uint32_t nw = wn;
            if (gn && gf) {
                if (sp - base == K) {  // ring full: spill the oldest entry
                    const uint32_t at = ((uint32_t)base % K) * WAVE + lane;
                    spill[base] = make_uint4(st_w[at], __float_as_uint(st_lo[at]), __float_as_uint(st_hi[at]), 0u);
                    ++base;
                }
                const uint32_t at = ((uint32_t)sp % K) * WAVE + lane;
                st_w[at] = wf;
                st_lo[at] = f_lo;
                st_hi[at] = hi;
                ++sp;
                hi = n_hi;
            } else if (gn) {
                hi = n_hi;
            } else if (gf) {
                nw = wf;
                lo = f_lo;
            } else {
                need_pop = true;
            }
            if (!need_pop) {
                w = nw;
                if ((w & 3u) == LEAF_TAG) {
                    state = ST_LEAF;
                    k = w >> 2;
                }
            }
        }
        // ---------------------------------------------------- leaf step --
        // (without leaf batching also for lanes that reached their leaf in
        // this iteration's node step: the leaf branch runs anyway while any
        // lane is in a leaf)
        if (run_leaf && !need_pop && state == ST_LEAF) {
            uint32_t ref[ILP];
            float4 ta[ILP], tb[ILP], tc[ILP];
#pragma unroll
            for (int j = 0; j < ILP; ++j) load_tri<LAYOUT>(s, recs, k + j, ref[j], ta[j], tb[j], tc[j]);
#pragma unroll
            for (int j = 0; j < ILP; ++j) hold(ta[j]), hold(tb[j]), hold(tc[j]);
            bool leaf_done = false, found = false;
#pragma unroll
            for (int j = 0; j < ILP; ++j) {
                if (!leaf_done && !found) {
                    ++cr;
                    const uint32_t tri = ref[j] & ~REF_LAST;
                    float t;
                    if (prd::mt_hit(ox, oy, oz, dx, dy, dz, ta[j].x, ta[j].y, ta[j].z, tb[j].x, tb[j].y, tb[j].z,
                                    tc[j].x, tc[j].y, tc[j].z, A.eps, t)) {
                        if (!ANY) {
                            if (t > 0.f && (t < best_t || (t == best_t && tri < best_i))) {
                                best_t = t;
                                best_i = tri;
                            }
                        } else {
                            const float t2 = t * t;
                            if (t > 0.f && (A.inclusive ? (t2 <= lim2) : (t2 < lim2))) {
                                best_t = t;
                                best_i = tri;
                                found = true;
                            }
                        }
                    }
                    if (ref[j] & REF_LAST) leaf_done = true;
                }
            }
            if (found) {
                sp = 0;
                base = 0;
                need_pop = true;  // stack is empty: the pop below finishes the ray
            } else if (leaf_done) {
                need_pop = true;
            } else {
                k += ILP;
            }
        }

        // ----------------------------------------------------------- pop --
        if (need_pop) {
            bool got = false;
            while (sp > 0) {
                --sp;
                uint32_t ew;
                float elo, ehi;
                if (sp >= base) {
                    const uint32_t at = ((uint32_t)sp % K) * WAVE + lane;
                    ew = st_w[at];
                    elo = st_lo[at];
                    ehi = st_hi[at];
                } else {
                    const uint4 e = spill[sp];
                    ew = e.x;
                    elo = __uint_as_float(e.y);
                    ehi = __uint_as_float(e.z);
                    base = sp;
                }
                const float cap = ANY ? ehi : fminf(ehi, best_t);
                if (elo <= cap) {
                    w = ew;
                    lo = elo;
                    hi = ehi;
                    got = true;
                    break;
                }
            }
            if (got) {
                if ((w & 3u) == LEAF_TAG) {
                    state = ST_LEAF;
                    k = w >> 2;
                } else {
                    state = ST_NODE;
                }
            } else {
                A.out_idx[ray] = (int32_t)best_i;
                if (A.out_t) A.out_t[ray] = best_i == NO_HIT ? FLT_MAX : best_t;
                if (A.counts) A.counts[2 * (size_t)ray] = cn, A.counts[2 * (size_t)ray + 1] = cr;
                state = ST_IDLE;
            }
        }
    }
}

// Two-level node records (SAH trees, whose children pairs have two equal
// clips).  For the node Z whose children pair is p (the pair a word w with
// w >> 3 == p names), quads[2p + i] describes Z's child c_i in 16 bytes:
//   .x  c_i's split plane (float bits), or its first leaf reference
//   .y  bits [29:0] the word of c_i's first child (30-bit form), [31:30] c_i's axis
//   .z  bits [29:0] the word of c_i's second child, [31:30] c_i's kind (Q_INNER,
//       Q_LEAF, Q_EMPTY)
//   .w  quads[2p].w: Z's split plane (float bits); quads[2p + 1].w unused
// A word in 30-bit form is the word itself, or Q_EMPTY30 for an empty leaf
// (indices stay below 2^28 - 1).  One 32-byte record, two loads from one cache
// line, carries a traversal two levels down (Z, then its near child); a far
// child pushed at the first level is a half reference (Q_HALF | p << 1 | i)
// whose 16 bytes alone carry one level.
constexpr uint32_t Q_INNER = 0u, Q_LEAF = 1u, Q_EMPTY = 2u;
constexpr uint32_t Q_EMPTY30 = 0x3FFFFFFFu;
constexpr uint32_t Q_MASK30 = 0x3FFFFFFFu;
constexpr uint32_t Q_HALF = 0x80000000u;

// Builds quads[2p], quads[2p + 1] for every pair p >= 1 (pair 0 holds the
// padding node and the root, which is no node's child).  flag bit 0: some pair
// has two different clips (not a SAH-form tree); bit 1: an index does not fit
// the 30-bit word form.  Either leaves the records unusable.
static __global__ void k_build_quads(const uint4 *pairs, uint32_t n_pairs, uint4 *quads, uint32_t *flag) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0 || p >= n_pairs) return;
    const uint4 pr = pairs[p];
    uint32_t bad = pr.x != pr.z ? 1u : 0u;
    uint4 h[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const uint32_t w = i == 0 ? pr.y : pr.w;
        uint4 q = make_uint4(0u, 0u, Q_EMPTY << 30, 0u);
        if (empty_leaf(w)) {
        } else if ((w & 3u) == LEAF_TAG) {
            q = make_uint4(w >> 2, 0u, Q_LEAF << 30, 0u);
            if ((w >> 2) >= (Q_MASK30 >> 2)) bad |= 2u;
        } else if ((w >> 3) < n_pairs) {
            const uint4 cp = pairs[w >> 3];
            uint32_t g[2] = {cp.y, cp.w};
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if (empty_leaf(g[j]))
                    g[j] = Q_EMPTY30;
                else if ((g[j] >> 2) >= (Q_MASK30 >> 2))  // keeps Q_EMPTY30 free
                    bad |= 2u;
            }
            q = make_uint4(cp.x, g[0] | ((w & 3u) << 30), g[1] | (Q_INNER << 30), 0u);
        } else {
            bad |= 2u;
        }
        h[i] = q;
    }
    h[0].w = pr.x;
    quads[2 * (size_t)p] = h[0];
    quads[2 * (size_t)p + 1] = h[1];
    if (bad) atomicOr(flag, bad);
}

// The closest-hit / any-hit traversal of k_trace (packed layout, single
// scene) over the two-level records: a node step loads one record and takes
// the node's level and, for the near (or only) child, that child's level,
// with the same plane arithmetic, widening, pruning and push order as two
// k_trace node steps, so every ray visits the same nodes and triangles and
// gets the same result and work counts.  Popping a far child of the first
// level loads its 16-byte half and takes its one level.  Each node step
// touches one cache line instead of one per level.
template <bool ANY, int K, int ILP, bool BATCH = false>
static __global__ __launch_bounds__(MAX_WG_WAVES * WAVE) void k_trace_q(DevScene s, TraceArgs A) {
    static_assert(K >= 2 && K <= 16, "LDS ring of 2..16 entries");
    extern __shared__ uint32_t ring_lds[];
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t wave_in_block = threadIdx.x / WAVE;
    const uint32_t gwave = blockIdx.x * (blockDim.x / WAVE) + wave_in_block;
    uint32_t *st_w = ring_lds + wave_in_block * (3 * K * WAVE);
    float *st_lo = reinterpret_cast<float *>(st_w + K * WAVE);
    float *st_hi = reinterpret_cast<float *>(st_w + 2 * K * WAVE);
    const uint64_t below = (1ull << lane) - 1ull;
    uint4 *spill = A.spill + ((size_t)gwave * WAVE + lane) * A.spill_depth;
    const uint32_t nn = A.n_dev ? *A.n_dev : A.n;
    const uint4 *pairs = reinterpret_cast<const uint4 *>(s.nodes);
    const uint4 *quads = s.quads;
    const float4 *recs = s.recs;
    // The root's word (node 1), read once per wave, not once per ray.
    const uint32_t root_w = (uint32_t)__builtin_amdgcn_readfirstlane((int)pairs[0].w);
    constexpr uint32_t ST_HALF = 4u;  // at a far child pushed by a record's first level

    uint32_t state = ST_IDLE, ray = 0, w = 0, k = 0, best_i = NO_HIT, cn = 0, cr = 0;
    float ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0, ix = 0, iy = 0, iz = 0, lim2 = 0;
    float lo = 0, hi = 0, best_t = INFINITY;
    int sp = 0, base = 0;
    const uint32_t nchunks = (nn + WAVE - 1) / WAVE;
    uint32_t pool_next = 0, pool_end = 0, ctr = gwave % NCTR;

    // Pushes (word, entry, exit) on the lane's stack, spilling the oldest LDS
    // ring entry to its HBM slot when the ring is full.
    auto push = [&](uint32_t pw, float plo, float phi) {
        if (sp - base == K) {
            const uint32_t at = ((uint32_t)base % K) * WAVE + lane;
            spill[base] = make_uint4(st_w[at], __float_as_uint(st_lo[at]), __float_as_uint(st_hi[at]), 0u);
            ++base;
        }
        const uint32_t at = ((uint32_t)sp % K) * WAVE + lane;
        st_w[at] = pw;
        st_lo[at] = plo;
        st_hi[at] = phi;
        ++sp;
    };

    while (true) {
        // ------------------------------------------------------- refill --
        const uint64_t idle = __ballot(state == ST_IDLE);
        const uint64_t dead = __ballot(state == ST_DEAD);
        if (dead == ~0ull) break;
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (nidle && (nidle >= A.fetch_min || (idle | dead) == ~0ull)) {
            const uint32_t avail = pool_end - pool_next;
            uint32_t nb = 0;
            if (avail < nidle) {
                const uint32_t first = (uint32_t)(__ffsll((unsigned long long)idle) - 1);
                uint32_t c = nchunks;
                if (lane == first) {
                    for (uint32_t tries = 0; tries < NCTR; ++tries) {
                        const uint32_t own = ctr < nchunks ? (nchunks - ctr + NCTR - 1) / NCTR : 0u;
                        uint32_t *p = A.work + ctr * CTR_STRIDE;
                        if (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < own) {
                            const uint32_t i = atomicAdd(p, 1u);
                            if (i < own) {
                                c = ctr + i * NCTR;
                                break;
                            }
                        }
                        ctr = ctr + 1 == NCTR ? 0u : ctr + 1;
                    }
                }
                c = (uint32_t)__builtin_amdgcn_readlane((int)c, (int)first);
                ctr = (uint32_t)__builtin_amdgcn_readlane((int)ctr, (int)first);
                nb = c * WAVE;
            }
            const uint32_t r = (uint32_t)__popcll(idle & below);
            if (state == ST_IDLE) {
                ray = r < avail ? pool_next + r : nb + (r - avail);
                if (ray >= nn) {
                    state = ST_DEAD;
                } else {
                    const float2 r0 = reinterpret_cast<const float2 *>(A.rays)[3 * (size_t)ray];
                    const float2 r1 = reinterpret_cast<const float2 *>(A.rays)[3 * (size_t)ray + 1];
                    const float2 r2 = reinterpret_cast<const float2 *>(A.rays)[3 * (size_t)ray + 2];
                    ox = r0.x, oy = r0.y, oz = r1.x, dx = r1.y, dy = r2.x, dz = r2.y;
                    ix = 1.0f / dx, iy = 1.0f / dy, iz = 1.0f / dz;
                    lo = 0.f, hi = INFINITY;
                    slab(s.box[0], s.box[3], ox, ix, lo, hi);
                    slab(s.box[1], s.box[4], oy, iy, lo, hi);
                    slab(s.box[2], s.box[5], oz, iz, lo, hi);
                    best_t = INFINITY, best_i = NO_HIT, sp = 0, base = 0, cn = 0, cr = 0;
                    bool live = true;
                    if (ANY) {
                        lim2 = A.limit2[ray];
                        if (!(lim2 >= 0.f))
                            live = false;
                        else
                            hi = fminf(hi, whi(sqrtf(lim2)));
                    }
                    if (live && lo <= hi) {
                        cn = 1;
                        w = root_w;
                        if ((w & 3u) != LEAF_TAG) {
                            state = ST_NODE;
                        } else if (!empty_leaf(w)) {
                            state = ST_LEAF;
                            k = w >> 2;
                        } else {
                            live = false;
                        }
                    } else {
                        live = false;
                    }
                    if (!live) {
                        A.out_idx[ray] = -1;
                        if (A.out_t) A.out_t[ray] = FLT_MAX;
                        if (A.counts) A.counts[2 * (size_t)ray] = cn, A.counts[2 * (size_t)ray + 1] = 0;
                        state = ST_IDLE;
                    }
                }
            }
            if (avail < nidle) {
                pool_next = nb + (nidle - avail);
                pool_end = nb + WAVE;
            } else {
                pool_next += nidle;
            }
            pool_next = (uint32_t)__builtin_amdgcn_readfirstlane((int)pool_next);
            pool_end = (uint32_t)__builtin_amdgcn_readfirstlane((int)pool_end);
        }

        bool run_node = true, run_leaf = true;
        if constexpr (BATCH) {
            const uint64_t at_leaf = __ballot(state == ST_LEAF);
            const uint64_t at_node = __ballot(state == ST_NODE || state == ST_HALF);
            run_leaf = at_leaf != 0ull && ((uint32_t)__popcll(at_leaf) >= A.leaf_batch || at_node == 0ull);
            run_node = !run_leaf;
        }
        bool need_pop = false;
        // ---------------------------------------------------- node step --
        if (run_node && (state == ST_NODE || state == ST_HALF)) {
            // The child whose level comes next: its half record.  Both
            // halves of the record are loaded for every lane in the step (a
            // far child's lane needs one of them), so that lanes at a node
            // and lanes at a far child wait for one round trip together.
            const uint32_t p = state == ST_NODE ? w >> 3 : (w & ~Q_HALF) >> 1;
            const uint4 h0 = quads[2 * (size_t)p];
            const uint4 h1 = quads[2 * (size_t)p + 1];
            hold(h0), hold(h1);
            uint4 h;
            bool go = true;
            if (state == ST_NODE) {
                // first level: node w, from its record's two halves
                const uint32_t axis = w & 3u;
                cn += 2;
                const float oa = axis == 0 ? ox : (axis == 1 ? oy : oz);
                const float ia = axis == 0 ? ix : (axis == 1 ? iy : iz);
                const float tz = (__uint_as_float(h0.w) - oa) * ia;
                const bool neg = ia < 0.f;
                const uint4 hn = neg ? h1 : h0, hf = neg ? h0 : h1;
                const float cap = ANY ? hi : fminf(hi, best_t);
                const float n_hi = fminf(cap, whi(tz));
                const float f_lo = fmaxf(lo, wlo(tz));
                const bool gn = lo <= n_hi && (hn.z >> 30) != Q_EMPTY;
                const bool gf = f_lo <= cap && (hf.z >> 30) != Q_EMPTY;
                h = hn;
                if (gn && gf) {
                    push(Q_HALF | (p << 1) | (neg ? 0u : 1u), f_lo, hi);
                    hi = n_hi;
                } else if (gn) {
                    hi = n_hi;
                } else if (gf) {
                    h = hf;
                    lo = f_lo;
                } else {
                    go = false;
                    need_pop = true;
                }
            } else {
                // a far child pushed by a first level: its half alone
                h = (w & 1u) ? h1 : h0;
            }
            if (go) {
                const uint32_t kind = h.z >> 30;
                if (kind == Q_LEAF) {
                    state = ST_LEAF;
                    k = h.x;
                } else {
                    // second level: the child itself
                    const uint32_t axis = h.y >> 30;
                    cn += 2;
                    const float oa = axis == 0 ? ox : (axis == 1 ? oy : oz);
                    const float ia = axis == 0 ? ix : (axis == 1 ? iy : iz);
                    const float tc = (__uint_as_float(h.x) - oa) * ia;
                    const bool neg = ia < 0.f;
                    const uint32_t w0 = h.y & Q_MASK30, w1 = h.z & Q_MASK30;
                    const uint32_t wn = neg ? w1 : w0, wf = neg ? w0 : w1;
                    const float cap = ANY ? hi : fminf(hi, best_t);
                    const float n_hi = fminf(cap, whi(tc));
                    const float f_lo = fmaxf(lo, wlo(tc));
                    const bool gn = lo <= n_hi && wn != Q_EMPTY30;
                    const bool gf = f_lo <= cap && wf != Q_EMPTY30;
                    uint32_t nw = wn;
                    if (gn && gf) {
                        push(wf, f_lo, hi);
                        hi = n_hi;
                    } else if (gn) {
                        hi = n_hi;
                    } else if (gf) {
                        nw = wf;
                        lo = f_lo;
                    } else {
                        need_pop = true;
                    }
                    if (!need_pop) {
                        w = nw;
                        if ((w & 3u) == LEAF_TAG) {
                            state = ST_LEAF;
                            k = w >> 2;
                        } else {
                            state = ST_NODE;
                        }
                    }
                }
            }
        }
        // ---------------------------------------------------- leaf step --
        if (run_leaf && !need_pop && state == ST_LEAF) {
            uint32_t ref[ILP];
            float4 ta[ILP], tb[ILP], tc[ILP];
#pragma unroll
            for (int j = 0; j < ILP; ++j) load_tri<1>(s, recs, k + j, ref[j], ta[j], tb[j], tc[j]);
#pragma unroll
            for (int j = 0; j < ILP; ++j) hold(ta[j]), hold(tb[j]), hold(tc[j]);
            bool leaf_done = false, found = false;
#pragma unroll
            for (int j = 0; j < ILP; ++j) {
                if (!leaf_done && !found) {
                    ++cr;
                    const uint32_t tri = ref[j] & ~REF_LAST;
                    float t;
                    if (prd::mt_hit(ox, oy, oz, dx, dy, dz, ta[j].x, ta[j].y, ta[j].z, tb[j].x, tb[j].y, tb[j].z,
                                    tc[j].x, tc[j].y, tc[j].z, A.eps, t)) {
                        if (!ANY) {
                            if (t > 0.f && (t < best_t || (t == best_t && tri < best_i))) {
                                best_t = t;
                                best_i = tri;
                            }
                        } else {
                            const float t2 = t * t;
                            if (t > 0.f && (A.inclusive ? (t2 <= lim2) : (t2 < lim2))) {
                                best_t = t;
                                best_i = tri;
                                found = true;
                            }
                        }
                    }
                    if (ref[j] & REF_LAST) leaf_done = true;
                }
            }
            if (found) {
                sp = 0;
                base = 0;
                need_pop = true;  // stack is empty: the pop below finishes the ray
            } else if (leaf_done) {
                need_pop = true;
            } else {
                k += ILP;
            }
        }

        // ----------------------------------------------------------- pop --
        if (need_pop) {
            bool got = false;
            while (sp > 0) {
                --sp;
                uint32_t ew;
                float elo, ehi;
                if (sp >= base) {
                    const uint32_t at = ((uint32_t)sp % K) * WAVE + lane;
                    ew = st_w[at];
                    elo = st_lo[at];
                    ehi = st_hi[at];
                } else {
                    const uint4 e = spill[sp];
                    ew = e.x;
                    elo = __uint_as_float(e.y);
                    ehi = __uint_as_float(e.z);
                    base = sp;
                }
                const float cap = ANY ? ehi : fminf(ehi, best_t);
                if (elo <= cap) {
                    w = ew;
                    lo = elo;
                    hi = ehi;
                    got = true;
                    break;
                }
            }
            if (got) {
                if (w & Q_HALF) {
                    state = ST_HALF;
                } else if ((w & 3u) == LEAF_TAG) {
                    state = ST_LEAF;
                    k = w >> 2;
                } else {
                    state = ST_NODE;
                }
            } else {
                A.out_idx[ray] = (int32_t)best_i;
                if (A.out_t) A.out_t[ray] = best_i == NO_HIT ? FLT_MAX : best_t;
                if (A.counts) A.counts[2 * (size_t)ray] = cn, A.counts[2 * (size_t)ray + 1] = cr;
                state = ST_IDLE;
            }
        }
    }
}

// LDS of one wave's ring: K entries of (word, lo, hi) per lane.
constexpr size_t ring_lds_bytes(int K) { return (size_t)3 * K * WAVE * sizeof(uint32_t); }

static __global__ void k_zero_u32(uint32_t *p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

static __global__ void k_set_u32(uint32_t *p, uint32_t v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

// Clears the NCTR chunk counters (one 64-thread block).
static __global__ void k_reset_counters(uint32_t *work) {
    if (threadIdx.x < NCTR) work[threadIdx.x * CTR_STRIDE] = 0u;
}

}  // namespace prt
