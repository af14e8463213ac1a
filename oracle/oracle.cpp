// ORACLE -- test infrastructure only.
//
// CPU restatement of the reference's tracing path (nicmue/nostrum_pray), used
// by tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py as
// the checker.  Nothing in the product (nostrum_pray_amd/, include/) links or
// calls this file.  Each function names the reference lines it follows.
// Compiled with -ffp-contract=off so expressions round exactly as written.
//
// Parity pins (see tests/test_oracle.py): scalar MT == AVX MT bitwise;
// closed-form MT and slab cases; BF == reference BFSearch compiled from
// source (oracle/_ref); exhaustive KD traversal == reference KDTree::intersect
// compiled from source; camera rays == reference calculatePrimRays compiled
// from source; Philox4x32-10 == the Random123 known-answer vectors; flat
// traversal twin == BF on every ray.  Functions with no reference
// counterpart (Philox, the polynomial sincos, the radiance-cache MLP) are
// pinned by their published definitions (known-answer vectors, series
// bounds, finite differences).
#include <immintrin.h>
#include <omp.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#define ORC_API extern "C" __attribute__((visibility("default")))

namespace {

constexpr float INFTY = std::numeric_limits<float>::max();
constexpr uint32_t NO_HIT = 0xFFFFFFFFu;

// ---------------------------------------------------------------- basics --
struct V {
    float x, y, z;
};
inline V sub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V add(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V mul(V a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V mulv(V a, V b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
// geometry.h:36 / :38
inline float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V cross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// geometry.h:28-34
inline void normalize(V &a) {
    const float n2 = a.x * a.x + a.y * a.y + a.z * a.z;
    if (n2 > 0) {
        const float inv = 1 / std::sqrt(n2);
        a.x *= inv, a.y *= inv, a.z *= inv;
    }
}
inline V ld(const float *p) { return {p[0], p[1], p[2]}; }

// Moeller-Trumbore, geometry.h:241-262 (scalar path).
inline bool mt(V o, V d, V a, V ab, V ac, float eps, float &t) {
    const V pvec = cross(d, ac);
    const float det = dot(ab, pvec);
    if (det < eps) return false;
    const float invDet = 1 / det;
    const V tvec = sub(o, a);
    const float u = dot(tvec, pvec) * invDet;
    if (u < 0.f || u > 1.f) return false;
    const V qvec = cross(tvec, ab);
    const float v = dot(d, qvec) * invDet;
    if (v < 0.f || u + v > 1.f) return false;
    t = dot(ac, qvec) * invDet;
    return true;
}

// Slab "line" test, geometry.h:151-166 (multiplies by 1/dir, std::min/max).
inline bool slabLine(const float *bmin, const float *bmax, V o, V inv) {
    const float t1x = (bmin[0] - o.x) * inv.x, t1y = (bmin[1] - o.y) * inv.y, t1z = (bmin[2] - o.z) * inv.z;
    const float t2x = (bmax[0] - o.x) * inv.x, t2y = (bmax[1] - o.y) * inv.y, t2z = (bmax[2] - o.z) * inv.z;
    const float nx = std::min(t1x, t2x), ny = std::min(t1y, t2y), nz = std::min(t1z, t2z);
    const float fx = std::max(t1x, t2x), fy = std::max(t1y, t2y), fz = std::max(t1z, t2z);
    const float tNear = std::max(std::max(nx, ny), std::max(nz, std::numeric_limits<float>::lowest()));
    const float tFar = std::min(std::min(fx, fy), std::min(fz, std::numeric_limits<float>::max()));
    return tNear <= tFar;
}

// --------------------------------------------------------------- AVX MT --
// geometry_simd.h:9-33 and geometry_simd.cpp:42-96.
inline void cross8(__m256 ax, __m256 ay, __m256 az, __m256 bx, __m256 by, __m256 bz, __m256 &rx, __m256 &ry,
                   __m256 &rz) {
    rx = _mm256_sub_ps(_mm256_mul_ps(ay, bz), _mm256_mul_ps(az, by));
    ry = _mm256_sub_ps(_mm256_mul_ps(az, bx), _mm256_mul_ps(ax, bz));
    rz = _mm256_sub_ps(_mm256_mul_ps(ax, by), _mm256_mul_ps(ay, bx));
}
inline __m256 dot8(__m256 ax, __m256 ay, __m256 az, __m256 bx, __m256 by, __m256 bz) {
    return _mm256_add_ps(_mm256_add_ps(_mm256_mul_ps(ax, bx), _mm256_mul_ps(ay, by)), _mm256_mul_ps(az, bz));
}

inline __m256 mt8(const float *a, const float *ab, const float *ac, const __m256 *o, const __m256 *d, __m256 mask,
                  float eps, __m256 &dist) {
    const __m256 abX = _mm256_set1_ps(ab[0]), abY = _mm256_set1_ps(ab[1]), abZ = _mm256_set1_ps(ab[2]);
    const __m256 acX = _mm256_set1_ps(ac[0]), acY = _mm256_set1_ps(ac[1]), acZ = _mm256_set1_ps(ac[2]);
    __m256 px, py, pz;
    cross8(d[0], d[1], d[2], acX, acY, acZ, px, py, pz);
    const __m256 det = dot8(abX, abY, abZ, px, py, pz);
    __m256 res = _mm256_and_ps(mask, _mm256_cmp_ps(det, _mm256_set1_ps(eps), _CMP_GE_OQ));
    if (_mm256_movemask_ps(res) == 0) return res;
    const __m256 invDet = _mm256_div_ps(_mm256_set1_ps(1.f), det);
    const __m256 tx = _mm256_sub_ps(o[0], _mm256_set1_ps(a[0]));
    const __m256 ty = _mm256_sub_ps(o[1], _mm256_set1_ps(a[1]));
    const __m256 tz = _mm256_sub_ps(o[2], _mm256_set1_ps(a[2]));
    const __m256 u = _mm256_mul_ps(dot8(tx, ty, tz, px, py, pz), invDet);
    res = _mm256_and_ps(res, _mm256_cmp_ps(u, _mm256_set1_ps(0.f), _CMP_GE_OQ));
    res = _mm256_and_ps(res, _mm256_cmp_ps(u, _mm256_set1_ps(1.f), _CMP_LE_OQ));
    if (_mm256_movemask_ps(res) == 0) return res;
    __m256 qx, qy, qz;
    cross8(tx, ty, tz, abX, abY, abZ, qx, qy, qz);
    const __m256 v = _mm256_mul_ps(dot8(d[0], d[1], d[2], qx, qy, qz), invDet);
    res = _mm256_and_ps(res, _mm256_cmp_ps(v, _mm256_set1_ps(0.f), _CMP_GE_OQ));
    res = _mm256_and_ps(res, _mm256_cmp_ps(_mm256_add_ps(u, v), _mm256_set1_ps(1.f), _CMP_LE_OQ));
    if (_mm256_movemask_ps(res) == 0) return res;
    dist = _mm256_mul_ps(dot8(acX, acY, acZ, qx, qy, qz), invDet);
    return res;
}

// Packet slab, geometry_simd.cpp:6-40 (divides by the direction).
inline __m256 slab8(const float *bmin, const float *bmax, const __m256 *o, const __m256 *d) {
    const __m256 t1x = _mm256_div_ps(_mm256_sub_ps(_mm256_set1_ps(bmin[0]), o[0]), d[0]);
    const __m256 t1y = _mm256_div_ps(_mm256_sub_ps(_mm256_set1_ps(bmin[1]), o[1]), d[1]);
    const __m256 t1z = _mm256_div_ps(_mm256_sub_ps(_mm256_set1_ps(bmin[2]), o[2]), d[2]);
    const __m256 t2x = _mm256_div_ps(_mm256_sub_ps(_mm256_set1_ps(bmax[0]), o[0]), d[0]);
    const __m256 t2y = _mm256_div_ps(_mm256_sub_ps(_mm256_set1_ps(bmax[1]), o[1]), d[1]);
    const __m256 t2z = _mm256_div_ps(_mm256_sub_ps(_mm256_set1_ps(bmax[2]), o[2]), d[2]);
    const __m256 tn = _mm256_max_ps(_mm256_max_ps(_mm256_min_ps(t1x, t2x), _mm256_min_ps(t1y, t2y)),
                                    _mm256_max_ps(_mm256_min_ps(t1z, t2z),
                                                  _mm256_set1_ps(std::numeric_limits<float>::lowest())));
    const __m256 tf = _mm256_min_ps(_mm256_min_ps(_mm256_max_ps(t1x, t2x), _mm256_max_ps(t1y, t2y)),
                                    _mm256_min_ps(_mm256_max_ps(t1z, t2z), _mm256_set1_ps(INFTY)));
    return _mm256_cmp_ps(tn, tf, _CMP_LE_OQ);
}

// ----------------------------------------------------------- Philox etc. --
// Philox4x32-10 (Salmon et al., SC 2011).
inline void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c[0] = n0, c[1] = n1, c[2] = n2, c[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}
inline float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// sin/cos of 2*pi*u: quadrant from 4u, Taylor polynomials on [0, pi/2).
inline void sincos2pi(float u, float &s, float &c) {
    const float x = u * 4.0f;
    const int q = (int)x;
    const float r = x - (float)q;
    const float a = r * 1.57079632679489661923f;
    const float a2 = a * a;
    float sp = -2.50521083854417188e-08f;
    sp = sp * a2 + 2.75573192239858907e-06f;
    sp = sp * a2 + -1.98412698412698413e-04f;
    sp = sp * a2 + 8.33333333333333333e-03f;
    sp = sp * a2 + -1.66666666666666667e-01f;
    const float sa = a + a * (a2 * sp);
    float cp = 2.08767569878680990e-09f;
    cp = cp * a2 + -2.75573192239858907e-07f;
    cp = cp * a2 + 2.48015873015873016e-05f;
    cp = cp * a2 + -1.38888888888888889e-03f;
    cp = cp * a2 + 4.16666666666666667e-02f;
    cp = cp * a2 + -0.5f;
    const float ca = 1.0f + a2 * cp;
    switch (q & 3) {
        case 0: s = sa, c = ca; break;
        case 1: s = ca, c = -sa; break;
        case 2: s = -sa, c = -ca; break;
        default: s = -ca, c = sa; break;
    }
}

// Cosine-weighted hemisphere sample, geometry.cpp:4-29, with
// sin(acos(sqrt(1-u1))) written as sqrt(u1); normalised twice (the second
// time by the Ray constructor, geometry.h:84-87).
inline V sampleHemisphere(V n, float u1, float u2) {
    V h = n;
    if (std::fabs(h.x) <= std::fabs(h.y) && std::fabs(h.x) <= std::fabs(h.z))
        h.x = 1.f;
    else if (std::fabs(h.y) <= std::fabs(h.x) && std::fabs(h.y) <= std::fabs(h.z))
        h.y = 1.f;
    else
        h.z = 1.f;
    V x = cross(h, n);
    normalize(x);
    V z = cross(x, n);
    normalize(z);
    const float st = std::sqrt(u1), ct = std::sqrt(1.0f - u1);
    float sp, cp;
    sincos2pi(u2, sp, cp);
    const V s = {st * cp, ct, st * sp};
    V d = add(add(mul(x, s.x), mul(n, s.y)), mul(z, s.z));
    normalize(d);
    normalize(d);
    return d;
}

// ---------------------------------------------------------- flat layout --
constexpr uint32_t LEAF_TAG = 3u, LEAF_EMPTY = 0x3FFFFFFFu, REF_LAST = 0x80000000u;

struct Flat {
    const uint32_t *nodes;
    const uint32_t *refs;
    const float *a, *ab, *ac;
    const float *box;
};

// Multiplicative widening: +-inf stays infinite (device twin).
inline float whi(float x) { return x * (x > 0.f ? 1.f + 1e-6f : 1.f - 1e-6f); }
inline float wlo(float x) { return x * (x > 0.f ? 1.f - 1e-6f : 1.f + 1e-6f); }

inline void slabW(float bmin, float bmax, float o, float inv, float &lo, float &hi) {
    float t0 = (bmin - o) * inv, t1 = (bmax - o) * inv;
    if (inv < 0.f) std::swap(t0, t1);
    lo = std::fmax(lo, wlo(t0));
    hi = std::fmin(hi, whi(t1));
}

// CPU twin of the GPU traversal (nostrum_pray_amd/csrc/pray_hip.hip,
// trace_one): same interval arithmetic, same visit order, same counters.
// mode 0 closest, 1 any (t*t < lim2), 2 any inclusive (t*t <= lim2).
uint32_t flatTrace(const Flat &f, V o, V d, int mode, float lim2, float eps, float &tout, uint32_t &cn, uint32_t &cr) {
    const V inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    float lo = 0.f, hi = INFINITY;
    slabW(f.box[0], f.box[3], o.x, inv.x, lo, hi);
    slabW(f.box[1], f.box[4], o.y, inv.y, lo, hi);
    slabW(f.box[2], f.box[5], o.z, inv.z, lo, hi);
    if (mode != 0) {
        if (!(lim2 >= 0.f)) return NO_HIT;
        hi = std::fmin(hi, whi(std::sqrt(lim2)));
    }
    float best_t = INFINITY;
    uint32_t best_i = NO_HIT;
    if (!(lo <= hi)) return best_i;
    struct E {
        uint32_t w;
        float lo, hi;
    };
    E st[64];
    int sp = 0;
    uint32_t w = f.nodes[3];
    cn += 1;
    const float oo[3] = {o.x, o.y, o.z}, ii[3] = {inv.x, inv.y, inv.z};
    while (true) {
        while ((w & 3u) != LEAF_TAG) {
            const uint32_t axis = w & 3u, lc = w >> 2;
            const uint32_t *pr = f.nodes + 2 * lc;
            cn += 2;
            float cl, cr_;
            std::memcpy(&cl, pr, 4);
            std::memcpy(&cr_, pr + 2, 4);
            const float tL = (cl - oo[axis]) * ii[axis];
            const float tR = (cr_ - oo[axis]) * ii[axis];
            const bool neg = ii[axis] < 0.f;
            const float tn = neg ? tR : tL, tf = neg ? tL : tR;
            const uint32_t wn = neg ? pr[3] : pr[1], wf = neg ? pr[1] : pr[3];
            const float cap = mode == 0 ? std::fmin(hi, best_t) : hi;
            const float n_hi = std::fmin(cap, whi(tn));
            const float f_lo = std::fmax(lo, wlo(tf));
            const bool gn = lo <= n_hi, gf = f_lo <= cap;
            if (gn && gf) {
                st[sp++] = {wf, f_lo, hi};
                w = wn;
                hi = n_hi;
            } else if (gn) {
                w = wn;
                hi = n_hi;
            } else if (gf) {
                w = wf;
                lo = f_lo;
            } else {
                w = (LEAF_EMPTY << 2) | LEAF_TAG;
            }
        }
        uint32_t k = w >> 2;
        if (k != LEAF_EMPTY) {
            while (true) {
                const uint32_t ref = f.refs[k];
                const uint32_t tri = ref & ~REF_LAST;
                cr += 1;
                float t;
                if (mt(o, d, ld(f.a + 3 * (size_t)tri), ld(f.ab + 3 * (size_t)tri), ld(f.ac + 3 * (size_t)tri), eps,
                       t)) {
                    if (mode == 0) {
                        if (t > 0.f && (t < best_t || (t == best_t && tri < best_i))) best_t = t, best_i = tri;
                    } else {
                        const bool occ = mode == 1 ? (t * t < lim2) : (t * t <= lim2);
                        if (t > 0.f && occ) {
                            tout = t;
                            return tri;
                        }
                    }
                }
                if (ref & REF_LAST) break;
                ++k;
            }
        }
        bool got = false;
        while (sp > 0) {
            const E e = st[--sp];
            const float cap = mode == 0 ? std::fmin(e.hi, best_t) : e.hi;
            if (e.lo <= cap) {
                w = e.w, lo = e.lo, hi = e.hi;
                got = true;
                break;
            }
        }
        if (!got) break;
    }
    tout = best_t;
    return best_i;
}

// Twin of the boxed node step over the boxed layout (pray_boxed_layout,
// node_levels 3): 16 words per pair, per axis {child 0 lower, child 1 lower,
// child 0 upper, child 1 upper}, then the children's words; pair 0's child 1
// is the root.  Each node step tests both children's boxes against the
// ray's interval capped at the best t, enters the child with the smaller
// entry distance (child 0 on a tie) and stacks the other; leaves, stack and
// counters as flatTrace.  The median tree (kd_tree.cpp:47-84) is an object
// partition whose nodes carry the box of their triangles, and the reference
// tests that box per node (kd_tree.cpp:142-279).
uint32_t boxedTrace(const uint32_t *rec, const Flat &f, V o, V d, int mode, float lim2, float eps, float &tout,
                    uint32_t &cn, uint32_t &cr) {
    const V inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    float lo = 0.f, hi = INFINITY;
    slabW(f.box[0], f.box[3], o.x, inv.x, lo, hi);
    slabW(f.box[1], f.box[4], o.y, inv.y, lo, hi);
    slabW(f.box[2], f.box[5], o.z, inv.z, lo, hi);
    if (mode != 0) {
        if (!(lim2 >= 0.f)) return NO_HIT;
        hi = std::fmin(hi, whi(std::sqrt(lim2)));
    }
    float best_t = INFINITY;
    uint32_t best_i = NO_HIT;
    if (!(lo <= hi)) return best_i;
    struct E {
        uint32_t w;
        float lo, hi;
    };
    E st[128];
    int sp = 0;
    uint32_t w = rec[13];
    cn += 1;
    auto fl = [](uint32_t u) {
        float x;
        std::memcpy(&x, &u, 4);
        return x;
    };
    while (true) {
        while ((w & 3u) != LEAF_TAG) {
            const uint32_t *r = rec + 16 * (size_t)(w >> 3);
            cn += 2;
            const float cap = mode == 0 ? std::fmin(hi, best_t) : hi;
            float l[2], h[2];
            bool g[2];
            for (int c = 0; c < 2; c++) {
                l[c] = lo, h[c] = cap;
                slabW(fl(r[c]), fl(r[2 + c]), o.x, inv.x, l[c], h[c]);
                slabW(fl(r[4 + c]), fl(r[6 + c]), o.y, inv.y, l[c], h[c]);
                slabW(fl(r[8 + c]), fl(r[10 + c]), o.z, inv.z, l[c], h[c]);
                g[c] = l[c] <= h[c] && r[12 + c] != ((LEAF_EMPTY << 2) | LEAF_TAG);
            }
            const int n = (g[1] && (!g[0] || l[1] < l[0])) ? 1 : 0;
            if (g[0] && g[1]) st[sp++] = {r[12 + 1 - n], l[1 - n], h[1 - n]};
            if (g[n]) {
                w = r[12 + n], lo = l[n], hi = h[n];
            } else {
                w = (LEAF_EMPTY << 2) | LEAF_TAG;
            }
        }
        uint32_t k = w >> 2;
        if (k != LEAF_EMPTY) {
            while (true) {
                const uint32_t ref = f.refs[k];
                const uint32_t tri = ref & ~REF_LAST;
                cr += 1;
                float t;
                if (mt(o, d, ld(f.a + 3 * (size_t)tri), ld(f.ab + 3 * (size_t)tri), ld(f.ac + 3 * (size_t)tri), eps,
                       t)) {
                    if (mode == 0) {
                        if (t > 0.f && (t < best_t || (t == best_t && tri < best_i))) best_t = t, best_i = tri;
                    } else {
                        const bool occ = mode == 1 ? (t * t < lim2) : (t * t <= lim2);
                        if (t > 0.f && occ) {
                            tout = t;
                            return tri;
                        }
                    }
                }
                if (ref & REF_LAST) break;
                ++k;
            }
        }
        bool got = false;
        while (sp > 0) {
            const E e = st[--sp];
            const float cap = mode == 0 ? std::fmin(e.hi, best_t) : e.hi;
            if (e.lo <= cap) {
                w = e.w, lo = e.lo, hi = e.hi;
                got = true;
                break;
            }
        }
        if (!got) break;
    }
    tout = best_t;
    return best_i;
}

// Twin of the 4-wide boxed node step (node_levels 4): record q is 32 words,
// {lower x of slots 0..3}, {upper x}, {lower y}, {upper y}, {lower z},
// {upper z}, {slot words}; a step intersects the ray's interval, capped at
// the best t, with the four slots' boxes, orders the entered ones by entry
// distance (lower slot first on a tie), enters the nearest and stacks the
// others farthest first; leaves, stack and counters as boxedTrace, with 4
// counted per step.
uint32_t boxedTrace4(const uint32_t *rec, uint32_t root, const Flat &f, V o, V d, int mode, float lim2, float eps,
                     float &tout, uint32_t &cn, uint32_t &cr) {
    const V inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    float lo = 0.f, hi = INFINITY;
    slabW(f.box[0], f.box[3], o.x, inv.x, lo, hi);
    slabW(f.box[1], f.box[4], o.y, inv.y, lo, hi);
    slabW(f.box[2], f.box[5], o.z, inv.z, lo, hi);
    if (mode != 0) {
        if (!(lim2 >= 0.f)) return NO_HIT;
        hi = std::fmin(hi, whi(std::sqrt(lim2)));
    }
    float best_t = INFINITY;
    uint32_t best_i = NO_HIT;
    if (!(lo <= hi)) return best_i;
    struct E {
        uint32_t w;
        float lo, hi;
    };
    std::vector<E> st;
    uint32_t w = root;
    cn += 1;
    auto fl = [](uint32_t u) {
        float x;
        std::memcpy(&x, &u, 4);
        return x;
    };
    const uint32_t EMPTY = (LEAF_EMPTY << 2) | LEAF_TAG;
    while (true) {
        while ((w & 3u) != LEAF_TAG) {
            const uint32_t *r = rec + 32 * (size_t)(w >> 3);
            cn += 4;
            const float cap = mode == 0 ? std::fmin(hi, best_t) : hi;
            float l[4], h[4];
            int order[4], n = 0;
            for (int j = 0; j < 4; j++) {
                l[j] = lo, h[j] = cap;
                slabW(fl(r[j]), fl(r[4 + j]), o.x, inv.x, l[j], h[j]);
                slabW(fl(r[8 + j]), fl(r[12 + j]), o.y, inv.y, l[j], h[j]);
                slabW(fl(r[16 + j]), fl(r[20 + j]), o.z, inv.z, l[j], h[j]);
                if (l[j] <= h[j] && r[24 + j] != EMPTY) order[n++] = j;
            }
            std::stable_sort(order, order + n, [&](int a, int b) { return l[a] < l[b]; });
            for (int q = n - 1; q >= 1; q--) st.push_back({r[24 + order[q]], l[order[q]], h[order[q]]});
            if (n) {
                w = r[24 + order[0]], lo = l[order[0]], hi = h[order[0]];
            } else {
                w = EMPTY;
            }
        }
        uint32_t k = w >> 2;
        if (k != LEAF_EMPTY) {
            while (true) {
                const uint32_t ref = f.refs[k];
                const uint32_t tri = ref & ~REF_LAST;
                cr += 1;
                float t;
                if (mt(o, d, ld(f.a + 3 * (size_t)tri), ld(f.ab + 3 * (size_t)tri), ld(f.ac + 3 * (size_t)tri), eps,
                       t)) {
                    if (mode == 0) {
                        if (t > 0.f && (t < best_t || (t == best_t && tri < best_i))) best_t = t, best_i = tri;
                    } else {
                        const bool occ = mode == 1 ? (t * t < lim2) : (t * t <= lim2);
                        if (t > 0.f && occ) {
                            tout = t;
                            return tri;
                        }
                    }
                }
                if (ref & REF_LAST) break;
                ++k;
            }
        }
        bool got = false;
        while (!st.empty()) {
            const E e = st.back();
            st.pop_back();
            const float cap = mode == 0 ? std::fmin(e.hi, best_t) : e.hi;
            if (e.lo <= cap) {
                w = e.w, lo = e.lo, hi = e.hi;
                got = true;
                break;
            }
        }
        if (!got) break;
    }
    tout = best_t;
    return best_i;
}

// ----------------------------------------------------- reference KD tree --
// A tree in depth-first preorder (left subtree before right): per node
// is_leaf, first/count into the leaf triangle list, box[6].
struct PreTree {
    const int32_t *ni;  // n x 4: is_leaf, axis, first, count
    const float *nf;    // n x 7: split, box[6]
    const uint32_t *tris;
    std::vector<uint32_t> next;  // index after each node's subtree
    uint64_t n;
    void build() {
        next.assign(n, 0);
        // Postorder sizes via an explicit stack of (node, phase).
        std::vector<std::pair<uint64_t, int>> st;
        if (n) st.push_back({0, 0});
        std::vector<uint64_t> rightOf(n, 0);
        while (!st.empty()) {
            auto &top = st.back();
            const uint64_t i = top.first;
            if (ni[4 * i] == 1) {
                next[i] = i + 1;
                st.pop_back();
                continue;
            }
            if (top.second == 0) {
                top.second = 1;
                st.push_back({i + 1, 0});
            } else if (top.second == 1) {
                top.second = 2;
                rightOf[i] = next[i + 1];
                st.push_back({next[i + 1], 0});
            } else {
                next[i] = next[rightOf[i]];
                st.pop_back();
            }
        }
    }
};

// KDTree::intersectRecursive scalar (kd_tree.cpp:237-279): visit every node
// whose box the line crosses, left then right; leaf-local strict minimum with
// t > 0, accepted when below the running distance.
bool refRec(const PreTree &T, uint64_t i, const float *A, const float *AB, const float *AC, V o, V d, V inv, float eps,
            uint32_t &triIdx, float &dist) {
    const float *box = T.nf + 7 * i + 1;
    if (!slabLine(box, box + 3, o, inv)) return false;
    if (T.ni[4 * i] == 0) {
        const bool hl = refRec(T, i + 1, A, AB, AC, o, d, inv, eps, triIdx, dist);
        const bool hr = refRec(T, T.next[i + 1], A, AB, AC, o, d, inv, eps, triIdx, dist);
        return hl || hr;
    }
    bool hit = false;
    float minDist = INFTY;
    uint32_t minIdx = 0;
    const uint32_t first = (uint32_t)T.ni[4 * i + 2], cnt = (uint32_t)T.ni[4 * i + 3];
    for (uint32_t k = 0; k < cnt; k++) {
        const uint32_t t = T.tris[first + k];
        float tmp;
        if (mt(o, d, ld(A + 3 * (size_t)t), ld(AB + 3 * (size_t)t), ld(AC + 3 * (size_t)t), eps, tmp)) {
            if (tmp < minDist && tmp > 0) {
                hit = true;
                minDist = tmp;
                minIdx = t;
            }
        }
    }
    if (hit && minDist < dist) {
        triIdx = minIdx;
        dist = minDist;
        return true;
    }
    return false;
}

// Packet version (kd_tree.cpp:282-367).
__m256 refRec8(const PreTree &T, uint64_t i, const float *A, const float *AB, const float *AC, const __m256 *o,
               const __m256 *d, __m256 mask, float eps, uint32_t *trsIdx, __m256 &distance) {
    const float *box = T.nf + 7 * i + 1;
    const __m256 bb = slab8(box, box + 3, o, d);
    if (_mm256_movemask_ps(bb) == 0) return bb;
    mask = _mm256_and_ps(mask, bb);
    if (T.ni[4 * i] == 0) {
        const __m256 hl = refRec8(T, i + 1, A, AB, AC, o, d, mask, eps, trsIdx, distance);
        const __m256 hr = refRec8(T, T.next[i + 1], A, AB, AC, o, d, mask, eps, trsIdx, distance);
        return _mm256_or_ps(hl, hr);
    }
    __m256 hit = _mm256_setzero_ps(), minDist = _mm256_set1_ps(INFTY);
    uint32_t minIdx[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t first = (uint32_t)T.ni[4 * i + 2], cnt = (uint32_t)T.ni[4 * i + 3];
    for (uint32_t k = 0; k < cnt; k++) {
        const uint32_t t = T.tris[first + k];
        __m256 tmp = _mm256_setzero_ps();
        const __m256 r = mt8(A + 3 * (size_t)t, AB + 3 * (size_t)t, AC + 3 * (size_t)t, o, d, mask, eps, tmp);
        const __m256 check = _mm256_and_ps(
            r, _mm256_and_ps(_mm256_cmp_ps(tmp, minDist, _CMP_LT_OQ), _mm256_cmp_ps(tmp, _mm256_setzero_ps(), _CMP_GT_OQ)));
        const int hm = _mm256_movemask_ps(check);
        if (hm == 0) continue;
        minDist = _mm256_blendv_ps(minDist, tmp, check);
        for (int j = 0; j < 8; j++)
            if ((hm >> j) & 1) minIdx[j] = t;
        hit = _mm256_or_ps(hit, check);
    }
    const __m256 endCheck = _mm256_and_ps(hit, _mm256_cmp_ps(minDist, distance, _CMP_LT_OQ));
    const int em = _mm256_movemask_ps(endCheck);
    for (int j = 0; j < 8; j++)
        if ((em >> j) & 1) trsIdx[j] = minIdx[j];
    distance = _mm256_blendv_ps(distance, minDist, endCheck);
    return endCheck;
}

void loadPacket(const float *rays, uint64_t n, uint64_t base, __m256 *o, __m256 *d) {
    alignas(32) float c[6][8];
    for (int j = 0; j < 8; j++) {
        const uint64_t r = base + j < n ? base + j : n - 1;  // pad with the last ray
        for (int k = 0; k < 6; k++) c[k][j] = rays[6 * r + k];
    }
    for (int k = 0; k < 3; k++) o[k] = _mm256_load_ps(c[k]), d[k] = _mm256_load_ps(c[3 + k]);
}

}  // namespace

// ================================================================== C ABI ==

ORC_API void orc_philox(const uint32_t *ctr, uint32_t k0, uint32_t k1, uint32_t *out) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    philox(c, k0, k1);
    std::memcpy(out, c, 16);
}

ORC_API void orc_sincos2pi(const float *u, uint64_t n, float *s, float *c) {
    for (uint64_t i = 0; i < n; i++) sincos2pi(u[i], s[i], c[i]);
}

ORC_API void orc_sample_hemisphere(const float *normals, const float *u1, const float *u2, uint64_t n, float *out) {
    for (uint64_t i = 0; i < n; i++) {
        const V d = sampleHemisphere(ld(normals + 3 * i), u1[i], u2[i]);
        out[3 * i] = d.x, out[3 * i + 1] = d.y, out[3 * i + 2] = d.z;
    }
}

// Ray stream of pray_gen_rays_device.
ORC_API void orc_gen_rays(uint64_t seed, const float *box6, uint64_t n, float *rays) {
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) {
        uint32_t r[4] = {(uint32_t)i, 0u, 0u, 0u}, q[4] = {(uint32_t)i, 1u, 0u, 0u};
        philox(r, k0, k1);
        philox(q, k0, k1);
        const float ox = box6[0] + box6[3] * u01(r[0]);
        const float oy = box6[1] + box6[4] * u01(r[1]);
        const float oz = box6[2] + box6[5] * u01(r[2]);
        const float z = 1.0f - 2.0f * u01(r[3]);
        const float rxy = std::sqrt(std::fmax(0.f, 1.0f - z * z));
        float sp, cp;
        sincos2pi(u01(q[0]), sp, cp);
        V dd = {rxy * cp, rxy * sp, z};
        normalize(dd);
        float *o = rays + 6 * i;
        o[0] = ox, o[1] = oy, o[2] = oz, o[3] = dd.x, o[4] = dd.y, o[5] = dd.z;
    }
}

// initialiseTriangles (scene_preprocessor.cpp:79-94): ab, ac, normal.
ORC_API void orc_init_tris(const float *abc, uint64_t n, float *ab, float *ac, float *normal) {
    for (uint64_t i = 0; i < n; i++) {
        const V a = ld(abc + 9 * i), b = ld(abc + 9 * i + 3), c = ld(abc + 9 * i + 6);
        const V e1 = sub(b, a), e2 = sub(c, a);
        V nn = cross(e1, e2);
        normalize(nn);
        std::memcpy(ab + 3 * i, &e1, 12);
        std::memcpy(ac + 3 * i, &e2, 12);
        std::memcpy(normal + 3 * i, &nn, 12);
    }
}

// Single MT test (scalar and AVX paths) for unit tests.
ORC_API int orc_mt(const float *ray6, const float *a, const float *ab, const float *ac, float eps, float *t) {
    float tt = 0;
    const bool h = mt(ld(ray6), ld(ray6 + 3), ld(a), ld(ab), ld(ac), eps, tt);
    *t = h ? tt : NAN;
    return h ? 1 : 0;
}

// AVX MT over rays (in packets of 8) against one triangle; hit mask + t.
ORC_API void orc_mt_avx(const float *rays, uint64_t n, const float *a, const float *ab, const float *ac, float eps,
                        int32_t *hit, float *t) {
    for (uint64_t b = 0; b < n; b += 8) {
        __m256 o[3], d[3];
        loadPacket(rays, n, b, o, d);
        __m256 dist = _mm256_set1_ps(NAN);
        const __m256 m = mt8(a, ab, ac, o, d, _mm256_castsi256_ps(_mm256_set1_epi32(-1)), eps, dist);
        alignas(32) float dv[8];
        _mm256_store_ps(dv, dist);
        const int mm = _mm256_movemask_ps(m);
        for (int j = 0; j < 8 && b + j < n; j++) {
            hit[b + j] = (mm >> j) & 1;
            t[b + j] = ((mm >> j) & 1) ? dv[j] : NAN;
        }
    }
}

// Reference line-slab test, scalar (geometry.h:151-166) and AVX
// (geometry_simd.cpp:6-40).
ORC_API void orc_slab(const float *rays, uint64_t n, const float *box6, int32_t *scalar, int32_t *avx) {
    for (uint64_t i = 0; i < n; i++) {
        const V o = ld(rays + 6 * i), d = ld(rays + 6 * i + 3);
        const V inv = {1 / d.x, 1 / d.y, 1 / d.z};
        scalar[i] = slabLine(box6, box6 + 3, o, inv) ? 1 : 0;
    }
    for (uint64_t b = 0; b < n; b += 8) {
        __m256 o[3], d[3];
        loadPacket(rays, n, b, o, d);
        const int m = _mm256_movemask_ps(slab8(box6, box6 + 3, o, d));
        for (int j = 0; j < 8 && b + j < n; j++) avx[b + j] = (m >> j) & 1;
    }
}

// Brute force (bf_search.cpp:26-40): global strict minimum over t > 0 in
// index order, so the lowest index wins at equal t.
ORC_API void orc_bf_closest(const float *A, const float *AB, const float *AC, uint64_t n_tris, const float *rays,
                            uint64_t n, float eps, int32_t *out_idx, float *out_t) {
#pragma omp parallel for schedule(dynamic, 16)
    for (uint64_t i = 0; i < n; i++) {
        const V o = ld(rays + 6 * i), d = ld(rays + 6 * i + 3);
        float dist = INFTY;
        uint32_t idx = NO_HIT;
        for (uint64_t k = 0; k < n_tris; k++) {
            float t;
            if (mt(o, d, ld(A + 3 * k), ld(AB + 3 * k), ld(AC + 3 * k), eps, t) && t > 0 && t < dist) {
                dist = t;
                idx = (uint32_t)k;
            }
        }
        out_idx[i] = (int32_t)idx;
        out_t[i] = dist;
    }
}

// Exhaustive reference traversal on a preorder tree export (scalar path).
ORC_API void orc_ref_closest(const int32_t *ni, const float *nf, uint64_t n_nodes, const uint32_t *tris,
                             const float *A, const float *AB, const float *AC, const float *rays, uint64_t n, float eps,
                             int32_t *out_idx, float *out_t) {
    PreTree T{ni, nf, tris, {}, n_nodes};
    T.build();
#pragma omp parallel for schedule(dynamic, 16)
    for (uint64_t i = 0; i < n; i++) {
        const V o = ld(rays + 6 * i), d = ld(rays + 6 * i + 3);
        const V inv = {1 / d.x, 1 / d.y, 1 / d.z};
        float dist = INFTY;
        uint32_t idx = 0;
        const bool hit = n_nodes && refRec(T, 0, A, AB, AC, o, d, inv, eps, idx, dist);
        out_idx[i] = hit ? (int32_t)idx : -1;
        out_t[i] = hit ? dist : INFTY;
    }
}

// Packet traversal timed as the CPU baseline: OpenMP schedule(dynamic,10)
// over 8-ray packets (pray.cpp:386) through the AVX recursion.  Returns the
// wall time in seconds.
ORC_API double orc_ref_closest_avx(const int32_t *ni, const float *nf, uint64_t n_nodes, const uint32_t *tris,
                                   const float *A, const float *AB, const float *AC, const float *rays, uint64_t n,
                                   float eps, int threads, int32_t *out_idx, float *out_t) {
    PreTree T{ni, nf, tris, {}, n_nodes};
    T.build();
    const uint64_t packets = (n + 7) / 8;
    const double t0 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 10) num_threads(threads)
    for (uint64_t p = 0; p < packets; p++) {
        __m256 o[3], d[3];
        loadPacket(rays, n, 8 * p, o, d);
        __m256 dist = _mm256_set1_ps(INFTY);
        uint32_t idx[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const __m256 h = n_nodes ? refRec8(T, 0, A, AB, AC, o, d, _mm256_castsi256_ps(_mm256_set1_epi32(-1)), eps, idx,
                                           dist)
                                 : _mm256_setzero_ps();
        alignas(32) float dv[8];
        _mm256_store_ps(dv, dist);
        const int m = _mm256_movemask_ps(h);
        for (int j = 0; j < 8 && 8 * p + j < n; j++) {
            out_idx[8 * p + j] = ((m >> j) & 1) ? (int32_t)idx[j] : -1;
            out_t[8 * p + j] = ((m >> j) & 1) ? dv[j] : INFTY;
        }
    }
    return omp_get_wtime() - t0;
}

// CPU twin of the GPU traversal over the flat layout.
ORC_API void orc_flat_trace(const uint32_t *nodes, const uint32_t *refs, const float *A, const float *AB,
                            const float *AC, const float *box6, uint64_t n_tris, const float *rays, uint64_t n,
                            int mode, float eps, const float *lim2, int32_t *out_idx, float *out_t, uint32_t *counts) {
    const Flat f{nodes, refs, A, AB, AC, box6};
#pragma omp parallel for schedule(dynamic, 64)
    for (uint64_t i = 0; i < n; i++) {
        uint32_t cn = 0, cr = 0;
        float t = INFINITY;
        uint32_t h = NO_HIT;
        if (n_tris) h = flatTrace(f, ld(rays + 6 * i), ld(rays + 6 * i + 3), mode, lim2 ? lim2[i] : 0.f, eps, t, cn, cr);
        out_idx[i] = (int32_t)h;
        out_t[i] = h == NO_HIT ? FLT_MAX : t;
        if (counts) counts[2 * i] = cn, counts[2 * i + 1] = cr;
    }
}

// flat_trace over the 4-wide boxed layout (boxedTrace4).
ORC_API void orc_trace_boxed4(const uint32_t *rec, uint32_t root, const uint32_t *brefs, const float *A,
                              const float *AB, const float *AC, const float *box6, uint64_t n_tris, const float *rays,
                              uint64_t n, int mode, float eps, const float *lim2, int32_t *out_idx, float *out_t,
                              uint32_t *counts) {
    const Flat f{nullptr, brefs, A, AB, AC, box6};
#pragma omp parallel for schedule(dynamic, 64)
    for (uint64_t i = 0; i < n; i++) {
        uint32_t cn = 0, cr = 0;
        float t = INFINITY;
        uint32_t h = NO_HIT;
        if (n_tris)
            h = boxedTrace4(rec, root, f, ld(rays + 6 * i), ld(rays + 6 * i + 3), mode, lim2 ? lim2[i] : 0.f, eps, t,
                            cn, cr);
        out_idx[i] = (int32_t)h;
        out_t[i] = h == NO_HIT ? FLT_MAX : t;
        if (counts) counts[2 * i] = cn, counts[2 * i + 1] = cr;
    }
}

// flat_trace over the boxed layout (boxedTrace): rec 16 words per pair,
// brefs its leaf runs (the last entry of each run flagged).
ORC_API void orc_trace_boxed(const uint32_t *rec, const uint32_t *brefs, const float *A, const float *AB,
                             const float *AC, const float *box6, uint64_t n_tris, const float *rays, uint64_t n,
                             int mode, float eps, const float *lim2, int32_t *out_idx, float *out_t,
                             uint32_t *counts) {
    const Flat f{nullptr, brefs, A, AB, AC, box6};
#pragma omp parallel for schedule(dynamic, 64)
    for (uint64_t i = 0; i < n; i++) {
        uint32_t cn = 0, cr = 0;
        float t = INFINITY;
        uint32_t h = NO_HIT;
        if (n_tris)
            h = boxedTrace(rec, f, ld(rays + 6 * i), ld(rays + 6 * i + 3), mode, lim2 ? lim2[i] : 0.f, eps, t, cn, cr);
        out_idx[i] = (int32_t)h;
        out_t[i] = h == NO_HIT ? FLT_MAX : t;
        if (counts) counts[2 * i] = cn, counts[2 * i + 1] = cr;
    }
}

// -------------------------------------------------------------- shading --
ORC_API void orc_camera_rays(const float *c, uint64_t w, uint64_t h, float *rays);

namespace {

struct Shade {
    Flat f;
    const float *normal, *color;
    const uint8_t *emit;
    float eps;
};

// Closest hit through the flat twin (equal to brute force on every ray).
inline uint32_t closest(const Shade &S, V o, V d, float &t) {
    uint32_t cn = 0, cr = 0;
    return flatTrace(S.f, o, d, 0, 0.f, S.eps, t, cn, cr);
}

// std::max(lo, std::min(hi, v)) with the standard library's argument order
// (geometry.h:15): NaN becomes hi.
inline float clampStd(float lo, float hi, float v) {
    const float m = (v < hi) ? v : hi;
    return (lo < m) ? m : lo;
}
inline uint8_t toByte(float c) { return static_cast<uint8_t>(c * 255.f + 0.5f); }

// Scalar evaluateRay (pray.cpp:17-55).
V evalScalar(const Shade &S, V o, V d, const float *lights, int nl, V bg) {
    float t;
    const uint32_t h = closest(S, o, d, t);
    if (h == NO_HIT) return bg;
    V color = {0.f, 0.f, 0.f};
    const V pos = add(o, mul(d, t));
    const V n = ld(S.normal + 3 * (size_t)h), alb = ld(S.color + 3 * (size_t)h);
    for (int l = 0; l < nl; l++) {
        const V lp = ld(lights + 6 * l), lc = ld(lights + 6 * l + 3);
        V L = sub(lp, pos);
        const float dist2 = L.x * L.x + L.y * L.y + L.z * L.z;
        normalize(L);
        const float lambert = dot(L, n);
        if (lambert < 0.f) continue;
        V Ld = L;
        normalize(Ld);  // Ray(intersectionPos, L) normalises again
        float st;
        const uint32_t sh = closest(S, pos, Ld, st);
        if (sh != NO_HIT && dist2 > st * st) continue;
        const float inv = 1 / dist2;
        color = add(color, mul(mul(mulv(alb, lc), lambert), inv));
        if (color.x >= 1.f && color.y >= 1.f && color.z >= 1.f) break;
    }
    return {clampStd(0.f, 1.f, color.x), clampStd(0.f, 1.f, color.y), clampStd(0.f, 1.f, color.z)};
}

// AVX evaluateRay for one 2x4 packet (pray.cpp:58-196), lanes in the
// reference's order (row y: x..x+3, row y+1: x..x+3).
void evalPacket(const Shade &S, const float *rays, const float *lights, int nl, V bg, const uint64_t pix[8], V out[8],
                bool &noHit) {
    V o[8], d[8], pos[8], n[8], alb[8], col[8];
    bool hit[8];
    float t[8];
    uint32_t tri[8];
    noHit = true;
    for (int j = 0; j < 8; j++) {
        o[j] = ld(rays + 6 * pix[j]);
        d[j] = ld(rays + 6 * pix[j] + 3);
        tri[j] = closest(S, o[j], d[j], t[j]);
        hit[j] = tri[j] != NO_HIT;
        noHit = noHit && !hit[j];
    }
    if (noHit) {
        for (int j = 0; j < 8; j++) out[j] = bg;
        return;
    }
    for (int j = 0; j < 8; j++) {
        col[j] = hit[j] ? V{0.f, 0.f, 0.f} : bg;
        pos[j] = add(o[j], mul(d[j], t[j]));
        n[j] = hit[j] ? ld(S.normal + 3 * (size_t)tri[j]) : V{0.f, 0.f, 0.f};
        alb[j] = hit[j] ? ld(S.color + 3 * (size_t)tri[j]) : V{0.f, 0.f, 0.f};
    }
    for (int l = 0; l < nl; l++) {
        const V lp = ld(lights + 6 * l), lc = ld(lights + 6 * l + 3);
        V L[8];
        float dist2[8], lambert[8];
        bool smask[8], any = false;
        for (int j = 0; j < 8; j++) {
            L[j] = sub(lp, pos[j]);
            dist2[j] = L[j].x * L[j].x + L[j].y * L[j].y + L[j].z * L[j].z;
            const float inv = 1.f / std::sqrt(dist2[j]);  // normalize_avx: no zero guard
            L[j] = mul(L[j], inv);
            lambert[j] = dot(L[j], n[j]);
            smask[j] = hit[j] && lambert[j] >= 0.f;
            any = any || smask[j];
        }
        if (!any) continue;
        bool lit[8], anyLit = false;
        for (int j = 0; j < 8; j++) {
            bool occluded = false;
            if (smask[j]) {
                float st;
                const uint32_t sh = closest(S, pos[j], L[j], st);
                occluded = sh != NO_HIT && dist2[j] >= st * st;
            }
            // the lambert mask is not re-applied here (pray.cpp:129)
            lit[j] = hit[j] && !occluded;
            anyLit = anyLit || lit[j];
        }
        if (!anyLit) continue;
        bool white = true;
        for (int j = 0; j < 8; j++) {
            if (lit[j]) {
                const float decay = lambert[j] / dist2[j];
                col[j].x = col[j].x + alb[j].x * (lc.x * decay);
                col[j].y = col[j].y + alb[j].y * (lc.y * decay);
                col[j].z = col[j].z + alb[j].z * (lc.z * decay);
            }
            white = white && col[j].x >= 1.f && col[j].y >= 1.f && col[j].z >= 1.f;
        }
        if (white) break;
    }
    for (int j = 0; j < 8; j++)
        out[j] = {clampStd(0.f, 1.f, col[j].x), clampStd(0.f, 1.f, col[j].y), clampStd(0.f, 1.f, col[j].z)};
}

}  // namespace

// rayTrace (pray.cpp:380-494): 2x4 AVX packets, then the scalar remainder
// columns (one extra column recomputed, quirk of pray.cpp:463-478) and an odd
// last row.  rgb: w*h*3 bytes, top row first.
ORC_API void orc_render_rt(const uint32_t *nodes, const uint32_t *refs, const float *A, const float *AB,
                           const float *AC, const float *box6, const float *normal, const float *color,
                           const float *cam16, uint64_t w, uint64_t h, const float *lights, int nl, const float *bg3,
                           float eps, uint8_t *rgb) {
    const Shade S{{nodes, refs, A, AB, AC, box6}, normal, color, nullptr, eps};
    std::vector<float> rays(6 * w * h);
    orc_camera_rays(cam16, w, h, rays.data());
    const V bg = ld(bg3);
    auto put = [&](uint64_t x, uint64_t y, V c) {
        uint8_t *p = rgb + 3 * (x + y * w);
        p[0] = toByte(c.x), p[1] = toByte(c.y), p[2] = toByte(c.z);
    };
    if (h >= 2 && w >= 4) {
        const uint64_t ty = (h - 2) / 2 + 1, tx = (w - 4) / 4 + 1;
#pragma omp parallel for schedule(dynamic, 10) collapse(2)
        for (uint64_t yt = 0; yt < ty; yt++) {
            for (uint64_t xt = 0; xt < tx; xt++) {
                const uint64_t x = 4 * xt, y = 2 * yt;
                uint64_t pix[8];
                for (int j = 0; j < 8; j++) pix[j] = (x + (j & 3)) + (y + (j >> 2)) * w;
                V out[8];
                bool noHit;
                evalPacket(S, rays.data(), lights, nl, bg, pix, out, noHit);
                for (int j = 0; j < 8; j++) put(x + (j & 3), y + (j >> 2), out[j]);
            }
        }
    }
    const uint64_t missing = w % 4;
    for (uint64_t y = 0; y + 1 < h; y++)
        for (uint64_t x = w - 1 - missing; x < w; x++)
            put(x, y, evalScalar(S, ld(&rays[6 * (x + y * w)]), ld(&rays[6 * (x + y * w) + 3]), lights, nl, bg));
    if (h % 2 != 0)
        for (uint64_t x = 0; x < w; x++)
            put(x, h - 1,
                evalScalar(S, ld(&rays[6 * (x + (h - 1) * w)]), ld(&rays[6 * (x + (h - 1) * w) + 3]), lights, nl, bg));
}

// Path tracing (NORMAL estimator, pray.cpp:342-378) with the build's
// wavefront conventions: throughput multiplied forward, Philox4x32-10 keyed
// by (seed_lo, seed_hi) with counter (pixel, sample, bounce, 0) giving
// u1, u2 (hemisphere) and u3 (Russian roulette from bounce rr_depth on:
// survive with q = max albedo component, throughput divided by q).
// Samples are summed in order; hdr = sum / spp; rgb = clamp -> byte.
// At max_depth <= 1 the reference takes its DIFFUSE estimator instead
// (scene_preprocessor.cpp:34-40, pray.cpp:315-340, evaluatePath :366-376):
// the camera ray is traced once; a miss gives the background and an
// emissive hit its emission; a non-emissive hit gives black at max_depth 0
// and otherwise, for each sample in order, the radiance arriving along a
// hemisphere ray of the same Philox counter (background, emission, or black
// at the depth limit) is summed, and the pixel is (sum * albedo) / spp.  No
// roulette.
ORC_API void orc_render_pt(const uint32_t *nodes, const uint32_t *refs, const float *A, const float *AB,
                           const float *AC, const float *box6, const float *normal, const float *color,
                           const uint8_t *emit, const float *cam16, uint64_t w, uint64_t h, uint32_t spp,
                           uint32_t max_depth, uint32_t rr_depth, uint64_t seed, const float *bg3, float eps,
                           float *hdr, uint8_t *rgb, uint64_t *queues) {
    // queues (optional, max_depth + 1 entries): paths traced per bounce over
    // all pixels and samples, the sizes of a wavefront tracer's queues.
    const Shade S{{nodes, refs, A, AB, AC, box6}, normal, color, emit, eps};
    if (queues)
        for (uint32_t b = 0; b <= max_depth; b++) queues[b] = 0;
    std::vector<float> rays(6 * w * h);
    orc_camera_rays(cam16, w, h, rays.data());
    const V bg = ld(bg3);
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    auto out = [&](uint64_t p, V m) {
        if (hdr) hdr[3 * p] = m.x, hdr[3 * p + 1] = m.y, hdr[3 * p + 2] = m.z;
        if (rgb) {
            rgb[3 * p] = toByte(clampStd(0.f, 1.f, m.x));
            rgb[3 * p + 1] = toByte(clampStd(0.f, 1.f, m.y));
            rgb[3 * p + 2] = toByte(clampStd(0.f, 1.f, m.z));
        }
    };
    if (max_depth <= 1) {
#pragma omp parallel for schedule(dynamic, 16)
        for (uint64_t p = 0; p < w * h; p++) {
            const V o = ld(&rays[6 * p]), d = ld(&rays[6 * p + 3]);
            if (queues) {
#pragma omp atomic
                queues[0] += spp;
            }
            float t;
            const uint32_t tri = closest(S, o, d, t);
            if (tri == NO_HIT) {
                out(p, bg);
                continue;
            }
            const V alb = ld(color + 3 * (size_t)tri);
            if (emit[tri]) {
                out(p, alb);
                continue;
            }
            if (max_depth == 0) {
                out(p, {0.f, 0.f, 0.f});
                continue;
            }
            if (queues) {
#pragma omp atomic
                queues[1] += spp;
            }
            const V pos = add(o, mul(d, t));
            const V n = ld(normal + 3 * (size_t)tri);
            V sum = {0.f, 0.f, 0.f};
            for (uint32_t s = 0; s < spp; s++) {
                uint32_t c[4] = {(uint32_t)p, s, 0u, 0u};
                philox(c, k0, k1);
                const V nd = sampleHemisphere(n, u01(c[0]), u01(c[1]));
                float t2;
                const uint32_t tri2 = closest(S, pos, nd, t2);
                // depth 1 == max_depth: only the background or an emitter add
                const V L = tri2 == NO_HIT ? bg : (emit[tri2] ? ld(color + 3 * (size_t)tri2) : V{0.f, 0.f, 0.f});
                sum = add(sum, L);
            }
            const float fs = (float)spp;
            out(p, {(sum.x * alb.x) / fs, (sum.y * alb.y) / fs, (sum.z * alb.z) / fs});
        }
        return;
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (uint64_t p = 0; p < w * h; p++) {
        V sum = {0.f, 0.f, 0.f};
        for (uint32_t s = 0; s < spp; s++) {
            V o = ld(&rays[6 * p]), d = ld(&rays[6 * p + 3]);
            V T = {1.f, 1.f, 1.f};
            V contrib = {0.f, 0.f, 0.f};
            for (uint32_t b = 0;; b++) {
                if (queues) {
#pragma omp atomic
                    queues[b]++;
                }
                float t;
                const uint32_t tri = closest(S, o, d, t);
                if (tri == NO_HIT) {
                    contrib = mulv(T, bg);
                    break;
                }
                if (emit[tri]) {
                    contrib = mulv(T, ld(color + 3 * (size_t)tri));
                    break;
                }
                if (b >= max_depth) break;
                const V alb = ld(color + 3 * (size_t)tri);
                uint32_t c[4] = {(uint32_t)p, s, b, 0u};
                philox(c, k0, k1);
                const V pos = add(o, mul(d, t));
                const V nd = sampleHemisphere(ld(normal + 3 * (size_t)tri), u01(c[0]), u01(c[1]));
                T = mulv(T, alb);
                if (b >= rr_depth) {
                    const float q = std::max(std::max(alb.x, alb.y), alb.z);
                    if (q < 1.f) {
                        if (!(u01(c[2]) < q)) break;
                        T = {T.x / q, T.y / q, T.z / q};
                    }
                }
                o = pos;
                d = nd;
            }
            sum = add(sum, contrib);
        }
        const float fs = (float)spp;
        out(p, {sum.x / fs, sum.y / fs, sum.z / fs});
    }
}

// Camera (scene_preprocessor.cpp:98-136): returns the 16 camera floats.
ORC_API void orc_camera(const float *pos, const float *look, uint64_t fov, uint64_t w, uint64_t h, float *cam16) {
    const V P = ld(pos), L = ld(look);
    const V target = add(P, L);
    const V up = {0, 1, 0};
    V zaxis = sub(target, P);
    normalize(zaxis);
    V xaxis = cross(up, zaxis);
    normalize(xaxis);
    const V yaxis = cross(zaxis, xaxis);
    float t[4][4] = {{xaxis.x, yaxis.x, zaxis.x, 0},
                     {xaxis.y, yaxis.y, zaxis.y, 0},
                     {xaxis.z, yaxis.z, zaxis.z, 0},
                     {-dot(xaxis, P), -dot(yaxis, P), -dot(zaxis, P), 1}};
    float s[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    // Matrix44::inverse (geometry.h:322-402).
    bool singular = false;
    for (int i = 0; i < 3 && !singular; i++) {
        int pivot = i;
        float ps = std::fabs(t[i][i]);
        for (int j = i + 1; j < 4; j++) {
            const float tmp = std::fabs(t[j][i]);
            if (tmp > ps) pivot = j, ps = tmp;
        }
        if (ps == 0) {
            singular = true;
            break;
        }
        if (pivot != i)
            for (int j = 0; j < 4; j++) std::swap(t[i][j], t[pivot][j]), std::swap(s[i][j], s[pivot][j]);
        for (int j = i + 1; j < 4; j++) {
            const float f = t[j][i] / t[i][i];
            for (int k = 0; k < 4; k++) t[j][k] -= f * t[i][k], s[j][k] -= f * s[i][k];
        }
    }
    for (int i = 3; i >= 0 && !singular; --i) {
        float f = t[i][i];
        if (f == 0) {
            singular = true;
            break;
        }
        for (int j = 0; j < 4; j++) t[i][j] /= f, s[i][j] /= f;
        for (int j = 0; j < i; j++) {
            f = t[j][i];
            for (int k = 0; k < 4; k++) t[j][k] -= f * t[i][k], s[j][k] -= f * s[i][k];
        }
    }
    if (singular)
        for (int r = 0; r < 4; r++)
            for (int q = 0; q < 4; q++) s[r][q] = r == q ? 1.f : 0.f;
    const float PI = static_cast<float>(std::acos(-1));
    cam16[0] = P.x, cam16[1] = P.y, cam16[2] = P.z;
    for (int r = 0; r < 3; r++)
        for (int q = 0; q < 3; q++) cam16[3 + 3 * r + q] = s[r][q];
    cam16[12] = std::tan(PI * 0.5f * fov / 180.f);
    cam16[13] = 1 / float(w);
    cam16[14] = 1 / float(h);
    cam16[15] = w / float(h);
}

// Primary rays from the 16 camera floats (scene_preprocessor.cpp:124-135).
ORC_API void orc_camera_rays(const float *c, uint64_t w, uint64_t h, float *rays) {
    for (uint64_t y = 0; y < h; y++) {
        for (uint64_t x = 0; x < w; x++) {
            const float px = (2 * ((x + 0.5f) * c[13]) - 1) * c[12];
            const float py = (1 - 2 * ((y + 0.5f) * c[14])) * c[12] / c[15];
            V d = {px * c[3] + py * c[6] + 1.f * c[9], px * c[4] + py * c[7] + 1.f * c[10],
                   px * c[5] + py * c[8] + 1.f * c[11]};
            normalize(d);
            float *o = rays + 6 * (x + y * w);
            o[0] = c[0], o[1] = c[1], o[2] = c[2], o[3] = d.x, o[4] = d.y, o[5] = d.z;
        }
    }
}

// ------------------------------------------------- radiance-cache MLP --
// CPU model of the cache network (no reference counterpart; pinned by
// central finite differences in double precision and by PyTorch autograd).
// Layers row-major [out][in]: 64x192, 4 x 64x64, 32x64; inputs are the
// encoding of 10 features (12 frequencies of sin/cos(2 pi 2^(k-1) x) for
// the 7 geometric features, albedo, a constant 1, zero padding) in the
// device's channel order (mlp_encode1).  With
// `emulate`, weights, encoded inputs, hidden activations and back-propagated
// gradients are rounded to bfloat16 (nearest even) where the device rounds
// them; sums are in double.  Without it everything is double.
namespace {

constexpr int MLP_IN = 192, MLP_W = 64, MLP_NP = 30720;
constexpr int MLP_OUT[6] = {64, 64, 64, 64, 64, 32};
constexpr int MLP_INS[6] = {192, 64, 64, 64, 64, 64};
constexpr int MLP_OFS[6] = {0, 12288, 16384, 20480, 24576, 28672};

inline float bf16r(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

// Channel c sits in slot s of lane group g of the 32-channel block m
// (c = 32m + 4g + s for s < 4, 32m + 16 + 4g + s - 4 otherwise); with
// q = 8m + s, j = q % 6, h = g / 2 it holds sin (g even) / cos (g odd) of
// 2 pi 2^(2j + h - 1) x_f, f = q / 6, for q < 42; albedo r, g, b and the
// constant 1 for q == 42; zero beyond.  The FP32 value is sincos2pi of the
// fraction of 2^(h-1) x_f followed by 2j angle doublings
// (sin 2a = 2 (sin a cos a), cos 2a = (cos a - sin a)(cos a + sin a)).
float mlp_encode1(const float *f, int c) {
    const int m = c >> 5, cc = c & 31;
    const int g = (cc & 15) >> 2, s = (cc & 3) + (cc >= 16 ? 4 : 0);
    const int q = 8 * m + s;
    if (q < 42) {
        const float x = std::ldexp(f[q / 6], (g >> 1) - 1);
        const float u = x - std::floor(x);
        float sn, cs;
        sincos2pi(u, sn, cs);
        for (int k = 0; k < 2 * (q % 6); k++) {
            const float s2 = 2.0f * (sn * cs);
            const float c2 = (cs - sn) * (cs + sn);
            sn = s2, cs = c2;
        }
        return (g & 1) ? cs : sn;
    }
    if (q == 42) return g < 3 ? f[7 + g] : 1.f;
    return 0.f;
}

struct MlpFwd {
    double x[MLP_IN];
    double h[5][MLP_W];  // post-ReLU (rounded when emulating)
    double y[3];
};

void mlp_forward1(const std::vector<double> &W, const float *f, bool emulate, MlpFwd &F) {
    for (int c = 0; c < MLP_IN; c++) {
        const float e = mlp_encode1(f, c);
        F.x[c] = emulate ? bf16r(e) : e;
    }
    const double *in = F.x;
    for (int L = 0; L < 5; L++) {
        for (int o = 0; o < MLP_W; o++) {
            double z = 0;
            const double *w = &W[MLP_OFS[L] + o * MLP_INS[L]];
            for (int k = 0; k < MLP_INS[L]; k++) z += w[k] * in[k];
            const double a = z > 0 ? z : 0;
            F.h[L][o] = emulate ? bf16r((float)a) : a;
        }
        in = F.h[L];
    }
    for (int o = 0; o < 3; o++) {
        double z = 0;
        const double *w = &W[MLP_OFS[5] + o * MLP_W];
        for (int k = 0; k < MLP_W; k++) z += w[k] * F.h[4][k];
        F.y[o] = z;
    }
}

std::vector<double> mlp_weights(const float *params, bool emulate) {
    std::vector<double> W(MLP_NP);
    for (int p = 0; p < MLP_NP; p++) W[p] = emulate ? bf16r(params[p]) : params[p];
    return W;
}

}  // namespace

ORC_API void orc_mlp_encode(const float *feats, uint64_t n, float *out) {
    for (uint64_t i = 0; i < n; i++)
        for (int c = 0; c < MLP_IN; c++) out[i * MLP_IN + c] = mlp_encode1(feats + 10 * i, c);
}

ORC_API void orc_mlp_forward(const float *params, const float *feats, uint64_t n, int emulate, double *out) {
    const std::vector<double> W = mlp_weights(params, emulate != 0);
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) {
        MlpFwd F;
        mlp_forward1(W, feats + 10 * i, emulate != 0, F);
        for (int r = 0; r < 3; r++) out[3 * i + r] = F.y[r];
    }
}

// Loss = mean over samples and channels of (y - t)^2 / D with D = y^2 + 0.01
// taken from this evaluation, or from `denom` (n x 3) when given; the
// gradient treats D as a constant.
ORC_API void orc_mlp_grad(const float *params, const float *feats, const float *targets, uint64_t n, int emulate,
                          const double *denom, double *grad, double *loss) {
    const bool em = emulate != 0;
    const std::vector<double> W = mlp_weights(params, em);
    const int nt = omp_get_max_threads();
    std::vector<std::vector<double>> G(nt, std::vector<double>(MLP_NP, 0.0));
    std::vector<double> Ls(nt, 0.0);
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) {
        std::vector<double> &g = G[omp_get_thread_num()];
        MlpFwd F;
        mlp_forward1(W, feats + 10 * i, em, F);
        double dz[MLP_W], dh[MLP_W];
        double dy[32] = {0};
        for (int r = 0; r < 3; r++) {
            const double y = F.y[r], t = targets[3 * i + r];
            const double D = denom ? denom[3 * i + r] : y * y + 0.01;
            Ls[omp_get_thread_num()] += (y - t) * (y - t) / D;
            const double d = 2.0 * (y - t) / D / (3.0 * (double)n);
            dy[r] = em ? bf16r((float)d) : d;
        }
        // output layer
        for (int o = 0; o < 3; o++)
            for (int k = 0; k < MLP_W; k++) g[MLP_OFS[5] + o * MLP_W + k] += dy[o] * F.h[4][k];
        for (int k = 0; k < MLP_W; k++) {
            double s = 0;
            for (int o = 0; o < 3; o++) s += W[MLP_OFS[5] + o * MLP_W + k] * dy[o];
            dh[k] = s;
        }
        for (int L = 4; L >= 0; L--) {
            for (int k = 0; k < MLP_W; k++) {
                const double v = F.h[L][k] > 0 ? dh[k] : 0.0;
                dz[k] = em ? bf16r((float)v) : v;
            }
            const double *in = L > 0 ? F.h[L - 1] : F.x;
            for (int o = 0; o < MLP_W; o++)
                for (int k = 0; k < MLP_INS[L]; k++) g[MLP_OFS[L] + o * MLP_INS[L] + k] += dz[o] * in[k];
            if (L == 0) break;
            for (int k = 0; k < MLP_W; k++) {
                double s = 0;
                for (int o = 0; o < MLP_W; o++) s += W[MLP_OFS[L] + o * MLP_W + k] * dz[o];
                dh[k] = s;
            }
        }
    }
    for (int p = 0; p < MLP_NP; p++) {
        double s = 0;
        for (int t = 0; t < nt; t++) s += G[t][p];
        grad[p] = s;
    }
    if (loss) {
        double s = 0;
        for (double x : Ls) s += x;
        *loss = s / (3.0 * (double)n);
    }
}

// One Adam step (beta1 0.9, beta2 0.99, eps 1e-8) on float master weights.
ORC_API void orc_adam(float *params, float *m, float *v, const double *grad, uint64_t n, uint32_t t, float lr) {
    for (uint64_t p = 0; p < n; p++) {
        const float g = (float)grad[p];
        m[p] = 0.9f * m[p] + (1.f - 0.9f) * g;
        v[p] = 0.99f * v[p] + (1.f - 0.99f) * g * g;
        const float mh = m[p] / (1.f - std::pow(0.9f, (float)t));
        const float vh = v[p] / (1.f - std::pow(0.99f, (float)t));
        params[p] = params[p] - lr * mh / (std::sqrt(vh) + 1e-8f);
    }
}

// ---------------------------------------------------------------------------
// GPU tree build twin (no reference counterpart: the reference builds on the
// CPU).  The bounding interval hierarchy over Morton-ordered triangles that
// pray_build_scene builds, restated top down: the range [f, l] of sorted
// positions splits after the last position that shares more than the range's
// common prefix of (code, position) with f.  The radix tree this recursion
// produces is the one Karras (2012) computes node by node, with the same
// numbering (a left child covering [f, g] is inner node g, a right child
// covering [g+1, l] is inner node g+1), so the device layout can be compared
// word for word.  Codes, boxes and splits follow the GPU kernel's arithmetic
// (30-bit codes of ((centre - lo) * inv) * 1024 per axis, x in bits 3k+2).
namespace lbvh {

struct Build {
    uint64_t n;
    uint32_t leaf;
    std::vector<float> lo, hi;        // triangle boxes
    std::vector<uint32_t> key, val;   // sorted codes and triangle indices
    std::vector<float> nlo, nhi;      // inner node boxes
    std::vector<uint32_t> cnt, first, last, height;
    std::vector<uint32_t> *nodes;     // 2 words per node
    std::vector<uint8_t> last_flag;
};

static int delta(const Build &B, int64_t i, int64_t j) {
    if (j < 0 || j >= (int64_t)B.n) return -1;
    const uint32_t a = B.key[i], b = B.key[j];
    if (a != b) return __builtin_clz(a ^ b);
    return 32 + __builtin_clz((uint32_t)i ^ (uint32_t)j);
}

static uint32_t spread3(uint32_t v) {
    uint32_t r = 0;
    for (int b = 0; b < 10; b++) r |= ((v >> b) & 1u) << (3 * b);
    return r;
}

// Box, count and height (with collapsed leaves) of the node covering
// [f, l]; `id` is its inner-node index when f < l.
static void rec(Build &B, uint64_t f, uint64_t l, uint64_t id, float blo[3], float bhi[3], uint32_t &count,
                uint32_t &h) {
    if (f == l) {
        const uint32_t t = B.val[f];
        for (int k = 0; k < 3; k++) blo[k] = B.lo[3 * t + k], bhi[k] = B.hi[3 * t + k];
        count = 1, h = 0;
        return;
    }
    const int common = delta(B, (int64_t)f, (int64_t)l);
    uint64_t g = f;
    while (g + 1 < l && delta(B, (int64_t)f, (int64_t)(g + 1)) > common) ++g;
    float llo[3], lhi[3], rlo[3], rhi[3];
    uint32_t lc, lh, rc, rh;
    rec(B, f, g, g, llo, lhi, lc, lh);
    rec(B, g + 1, l, g + 1, rlo, rhi, rc, rh);
    for (int k = 0; k < 3; k++) blo[k] = std::min(llo[k], rlo[k]), bhi[k] = std::max(lhi[k], rhi[k]);
    count = lc + rc;
    const uint32_t hl = (f == g || lc <= B.leaf) ? 0u : lh, hr = (g + 1 == l || rc <= B.leaf) ? 0u : rh;
    h = count <= B.leaf ? 0u : 1u + std::max(hl, hr);
    for (int k = 0; k < 3; k++) B.nlo[3 * id + k] = blo[k], B.nhi[3 * id + k] = bhi[k];
    B.cnt[id] = count, B.first[id] = (uint32_t)f, B.last[id] = (uint32_t)l, B.height[id] = h;
}

static uint32_t axis_of(const Build &B, uint64_t id) {
    const uint32_t x = B.key[B.first[id]] ^ B.key[B.last[id]];
    if (x) {
        const uint32_t m = (31u - (uint32_t)__builtin_clz(x)) % 3u;
        return m == 2u ? 0u : (m == 1u ? 1u : 2u);
    }
    const float ex = B.nhi[3 * id] - B.nlo[3 * id], ey = B.nhi[3 * id + 1] - B.nlo[3 * id + 1],
                ez = B.nhi[3 * id + 2] - B.nlo[3 * id + 2];
    return (ex >= ey && ex >= ez) ? 0u : (ey >= ez ? 1u : 2u);
}

// Node entries of the range [f, l] (inner node `id` when f < l) as a child
// of a node split along `axis`: {bound, word}.
static void entry(Build &B, uint64_t f, uint64_t l, uint64_t id, uint32_t axis, bool upper, uint32_t *out) {
    float b;
    if (f == l) {
        const uint32_t t = B.val[f];
        b = upper ? B.hi[3 * t + axis] : B.lo[3 * t + axis];
    } else {
        b = upper ? B.nhi[3 * id + axis] : B.nlo[3 * id + axis];
    }
    std::memcpy(&out[0], &b, 4);
    if (f == l || B.cnt[id] <= B.leaf) {
        B.last_flag[l] = 1;
        out[1] = ((uint32_t)f << 2) | 3u;
    } else {
        out[1] = ((uint32_t)(2 * (id + 1)) << 2) | axis_of(B, id);
    }
}

static void write(Build &B, uint64_t f, uint64_t l, uint64_t id) {
    if (f == l || B.cnt[id] <= B.leaf) return;  // a leaf: written by its parent
    const int common = delta(B, (int64_t)f, (int64_t)l);
    uint64_t g = f;
    while (g + 1 < l && delta(B, (int64_t)f, (int64_t)(g + 1)) > common) ++g;
    const uint32_t axis = axis_of(B, id);
    uint32_t *pair = B.nodes->data() + 4 * (id + 1);
    entry(B, f, g, g, axis, true, pair);
    entry(B, g + 1, l, g + 1, axis, false, pair + 2);
    write(B, f, g, g);
    write(B, g + 1, l, g + 1);
}

}  // namespace lbvh

// nodes: 2n entries of 2 words (4n words, zero where unused); refs: n;
// box: 6; returns the tree depth.
ORC_API uint32_t orc_lbvh_build(const float *abc, uint64_t n, uint32_t leaf_size, uint32_t *nodes, uint32_t *refs,
                                float *box) {
    lbvh::Build B;
    B.n = n, B.leaf = leaf_size;
    std::vector<uint32_t> words(4 * std::max<uint64_t>(n, 1), 0u);
    B.nodes = &words;
    B.lo.resize(3 * n), B.hi.resize(3 * n);
    for (uint64_t i = 0; i < n; i++) {
        const float *p = abc + 9 * i;
        for (int k = 0; k < 3; k++) {
            const float x = p[k], y = p[3 + k], z = p[6 + k];
            B.lo[3 * i + k] = std::min(x, std::min(y, z));
            B.hi[3 * i + k] = std::max(x, std::max(y, z));
        }
    }
    float bl[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, bh[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint64_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) bl[k] = std::min(bl[k], B.lo[3 * i + k]), bh[k] = std::max(bh[k], B.hi[3 * i + k]);
    for (int k = 0; k < 3; k++) box[k] = bl[k], box[3 + k] = bh[k];
    std::vector<uint32_t> code(n);
    for (uint64_t i = 0; i < n; i++) {
        uint32_t q[3];
        for (int k = 0; k < 3; k++) {
            const float ext = bh[k] - bl[k];
            const float inv = ext > 0.f ? 1.0f / ext : 0.f;
            const float c = (B.lo[3 * i + k] + B.hi[3 * i + k]) * 0.5f;
            const float u = ((c - bl[k]) * inv) * 1024.0f;
            q[k] = (uint32_t)std::fmax(0.f, std::fmin(1023.f, u));
        }
        code[i] = (lbvh::spread3(q[0]) << 2) | (lbvh::spread3(q[1]) << 1) | lbvh::spread3(q[2]);
    }
    B.val.resize(n);
    for (uint64_t i = 0; i < n; i++) B.val[i] = (uint32_t)i;
    std::stable_sort(B.val.begin(), B.val.end(), [&](uint32_t x, uint32_t y) { return code[x] < code[y]; });
    B.key.resize(n);
    for (uint64_t i = 0; i < n; i++) B.key[i] = code[B.val[i]];
    const uint64_t inner = n > 1 ? n - 1 : 1;
    B.nlo.assign(3 * inner, 0.f), B.nhi.assign(3 * inner, 0.f);
    B.cnt.assign(inner, 0), B.first.assign(inner, 0), B.last.assign(inner, 0), B.height.assign(inner, 0);
    B.last_flag.assign(std::max<uint64_t>(n, 1), 0);
    uint32_t depth = 0;
    if (n >= 2) {
        float rlo[3], rhi[3];
        uint32_t rc, rh;
        lbvh::rec(B, 0, n - 1, 0, rlo, rhi, rc, rh);
        depth = B.cnt[0] <= leaf_size ? 0u : B.height[0];
        if (B.cnt[0] <= leaf_size) {
            words[3] = 3u;
            B.last_flag[n - 1] = 1;
        } else {
            words[3] = (2u << 2) | lbvh::axis_of(B, 0);
            lbvh::write(B, 0, n - 1, 0);
        }
    } else if (n == 1) {
        words[3] = 3u;
        B.last_flag[0] = 1;
    }
    std::memcpy(nodes, words.data(), 4 * std::max<uint64_t>(n, 1) * 4);
    for (uint64_t k = 0; k < n; k++) refs[k] = B.val[k] | (B.last_flag[k] ? 0x80000000u : 0u);
    return depth;
}

// Twin of the GPU binned-SAH KD build (PRAY_BUILD_SAH): level by level, the
// same bins (64 per axis over the node's cell, clipped reference boxes), the
// same cost (traversal + intersection x area-weighted counts, 0.8 for an
// empty side, kd_tree_sah.cpp:141-166) evaluated in the same order, the same
// classification and stable partitions, and the same breadth-first numbering
// of children pairs and leaf runs.
namespace sahbuild {

struct Node {
    float lo[3], hi[3];
    uint32_t begin, count, entry;
    float plane;
};

static int bin(float x, float lo, float scale) {
    return (int)std::fmax(0.f, std::fmin(63.f, std::floor((x - lo) * scale)));
}

static float half_area(float dx, float dy, float dz) { return dx * dy + dy * dz + dz * dx; }

}  // namespace sahbuild

// nodes: entry words (2 per entry), grown as needed; returns the number of
// entries; refs: leaf runs; depth and box out.
ORC_API uint64_t orc_sah_build(const float *abc, uint64_t n, uint32_t leaf_size, uint32_t max_depth, float ct,
                               float ci, uint32_t *nodes_out, uint64_t nodes_cap, uint32_t *refs_out,
                               uint64_t refs_cap, uint64_t *n_refs, uint32_t *depth, float *box) {
    using sahbuild::Node;
    std::vector<float> lo(3 * n), hi(3 * n);
    for (uint64_t i = 0; i < n; i++) {
        const float *p = abc + 9 * i;
        for (int k = 0; k < 3; k++) {
            const float x = p[k], y = p[3 + k], z = p[6 + k];
            lo[3 * i + k] = std::min(x, std::min(y, z));
            hi[3 * i + k] = std::max(x, std::max(y, z));
        }
    }
    float bl[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, bh[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint64_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) bl[k] = std::min(bl[k], lo[3 * i + k]), bh[k] = std::max(bh[k], hi[3 * i + k]);
    for (int k = 0; k < 3; k++) box[k] = bl[k], box[3 + k] = bh[k];
    std::vector<uint32_t> entries(4, 0u), leaf;
    std::vector<Node> cur(1);
    for (int k = 0; k < 3; k++) cur[0].lo[k] = bl[k], cur[0].hi[k] = bh[k];
    cur[0].begin = 0, cur[0].count = (uint32_t)n, cur[0].entry = 1, cur[0].plane = 0.f;
    std::vector<uint32_t> refs(n);
    for (uint64_t i = 0; i < n; i++) refs[i] = (uint32_t)i;
    uint32_t pair_base = 1;
    *depth = 0;
    for (uint32_t level = 0;; level++) {
        std::vector<Node> next;
        std::vector<uint32_t> next_refs;
        uint32_t n_split = 0;
        for (const Node &N : cur) {
            int best_a = -1;
            float best_p = 0.f;
            const float dx = N.hi[0] - N.lo[0], dy = N.hi[1] - N.lo[1], dz = N.hi[2] - N.lo[2];
            const float sa = sahbuild::half_area(dx, dy, dz);
            if (N.count > leaf_size && level < max_depth && sa > 0.f) {
                uint32_t H[3][2][64] = {};
                for (uint32_t i = N.begin; i < N.begin + N.count; i++) {
                    const uint32_t t = refs[i];
                    for (int a = 0; a < 3; a++) {
                        const float ext = N.hi[a] - N.lo[a];
                        const float scale = ext > 0.f ? 64.0f / ext : 0.f;
                        const float clo = std::fmax(lo[3 * t + a], N.lo[a]);
                        const float chi = std::fmin(hi[3 * t + a], N.hi[a]);
                        H[a][0][sahbuild::bin(clo, N.lo[a], scale)]++;
                        H[a][1][sahbuild::bin(chi, N.lo[a], scale)]++;
                    }
                }
                float best = ci * (float)N.count;
                for (int a = 0; a < 3; a++) {
                    const float ext = N.hi[a] - N.lo[a];
                    if (!(ext > 0.f)) continue;
                    uint32_t nl = 0, ends = 0;
                    for (int b = 1; b < 64; b++) {
                        nl += H[a][0][b - 1];
                        ends += H[a][1][b - 1];
                        const uint32_t nr = N.count - ends;
                        const float p = N.lo[a] + ext * ((float)b * (1.0f / 64.0f));
                        float l[3] = {dx, dy, dz}, r[3] = {dx, dy, dz};
                        l[a] = p - N.lo[a];
                        r[a] = N.hi[a] - p;
                        const float pl = sahbuild::half_area(l[0], l[1], l[2]) / sa;
                        const float pr = sahbuild::half_area(r[0], r[1], r[2]) / sa;
                        float cost = ct + ci * (pl * (float)nl + pr * (float)nr);
                        if ((nl == 0 || nr == 0) && !(pl == 1.f || pr == 1.f)) cost *= 0.8f;
                        if (cost < best) best = cost, best_a = a, best_p = p;
                    }
                }
            }
            uint32_t *e = entries.data();
            (void)e;
            if (best_a < 0) {
                const uint32_t w = N.count ? (((uint32_t)leaf.size() << 2) | 3u) : ((0x3FFFFFFFu << 2) | 3u);
                entries[2 * N.entry] = 0, entries[2 * N.entry + 1] = w;
                std::memcpy(&entries[2 * N.entry], &N.plane, 4);
                for (uint32_t i = 0; i < N.count; i++)
                    leaf.push_back(refs[N.begin + i] | (i + 1 == N.count ? 0x80000000u : 0u));
                continue;
            }
            const uint32_t q = pair_base + n_split++;
            if (entries.size() < 4 * ((size_t)q + 1)) entries.resize(4 * ((size_t)q + 1), 0u);
            std::memcpy(&entries[2 * N.entry], &N.plane, 4);
            entries[2 * N.entry + 1] = ((2u * q) << 2) | (uint32_t)best_a;
            Node L = N, R = N;
            L.hi[best_a] = best_p, R.lo[best_a] = best_p;
            L.entry = 2 * q, R.entry = 2 * q + 1, L.plane = R.plane = best_p;
            std::vector<uint32_t> lr, rr;
            for (uint32_t i = N.begin; i < N.begin + N.count; i++) {
                const uint32_t t = refs[i];
                const float clo = std::fmax(lo[3 * t + best_a], N.lo[best_a]);
                const float chi = std::fmin(hi[3 * t + best_a], N.hi[best_a]);
                if (clo < best_p || (clo == best_p && chi == best_p)) lr.push_back(t);
                if (chi > best_p) rr.push_back(t);
            }
            L.begin = (uint32_t)next_refs.size(), L.count = (uint32_t)lr.size();
            next_refs.insert(next_refs.end(), lr.begin(), lr.end());
            R.begin = (uint32_t)next_refs.size(), R.count = (uint32_t)rr.size();
            next_refs.insert(next_refs.end(), rr.begin(), rr.end());
            next.push_back(L);
            next.push_back(R);
        }
        pair_base += n_split;
        if (n_split == 0) {
            *depth = level;
            break;
        }
        cur.swap(next);
        refs.swap(next_refs);
    }
    const uint64_t n_entries = 2ull * pair_base;
    entries.resize(2 * n_entries, 0u);
    if (nodes_out && nodes_cap >= 2 * n_entries) std::memcpy(nodes_out, entries.data(), 2 * n_entries * 4);
    if (refs_out && refs_cap >= leaf.size()) std::memcpy(refs_out, leaf.data(), leaf.size() * 4);
    *n_refs = leaf.size();
    return n_entries;
}
