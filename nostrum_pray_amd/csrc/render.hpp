// Renderers on top of the traversal kernel.
//
// Ray tracing (reference rayTrace, pray.cpp:380-494).  The reference shades
// 2x4 pixel packets with AVX (pray.cpp:58-196) and then re-renders the last
// column(s) and an odd last row with the scalar path (pray.cpp:17-55).  The
// two paths differ in rounding order, in the shadow comparison (>= vs >), in
// re-normalising the shadow direction, and in packet-wide behaviour: after
// the shadow query the lambert mask is not applied again, so a back-facing
// light adds a negative term whenever another lane of the packet faces it,
// and a packet stops at the first light after which all eight lanes are white.
// Here each packet is eight adjacent lanes of a wave (64 = 8 packets) and the
// packet-wide decisions are ballots over those eight lanes; scalar-path
// pixels get their own lanes after the packets.  Shadow rays of both paths go
// through one inclusive any-hit launch (a strict limit L2 is passed as the
// largest float below L2).
//
// Path tracing (NORMAL estimator, pray.cpp:342-378; path_trace.cu:395-464).
// Wavefront: per sample pass, generate fills queue 0 with one path per pixel;
// each bounce traces the queue (closest hit, device-side count) and a shade
// kernel terminates paths (miss -> throughput * background, emitter ->
// throughput * Le, depth limit or Russian roulette -> 0) or pushes the bounced
// ray to the other queue (one atomic per wave).  Each path writes its
// contribution exactly once per pass; an accumulate kernel adds the pass to
// the sum in sample order, so results do not depend on queue order.  Random
// numbers: Philox4x32-10 keyed by the 64-bit seed, counter (pixel, sample,
// bounce, 0); the sample index lives in device memory so one captured graph
// serves every pass.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_math.hpp"
#include "traverse.hpp"

namespace prr {

using prt::DevScene;
using prt::WAVE;

struct Cam {
    float px, py, pz;
    float m[9];
    float angle, invW, invH, aspect;
};

// Primary ray direction of the reference camera (scene_preprocessor.cpp:128-133).
__device__ __forceinline__ void camera_dir(const Cam &c, uint32_t xp, uint32_t yp, float &dx, float &dy, float &dz) {
    const float x = (2.0f * (((float)xp + 0.5f) * c.invW) - 1.0f) * c.angle;
    const float y = (1.0f - 2.0f * (((float)yp + 0.5f) * c.invH)) * c.angle / c.aspect;
    const float z = 1.0f;
    dx = x * c.m[0] + y * c.m[3] + z * c.m[6];
    dy = x * c.m[1] + y * c.m[4] + z * c.m[7];
    dz = x * c.m[2] + y * c.m[5] + z * c.m[8];
    prd::normalize3(dx, dy, dz);
}

// std::max(lo, std::min(hi, v)) (geometry.h:15): NaN maps to hi.
__device__ __forceinline__ float clamp_std(float v) {
    const float m = (v < 1.0f) ? v : 1.0f;
    return (0.0f < m) ? m : 0.0f;
}
__device__ __forceinline__ uint8_t to_byte(float c) { return (uint8_t)(uint32_t)(c * 255.f + 0.5f); }

// Pixel layout of the RT launch.  Lanes [0, n_avx) are packet lanes (8 per
// packet, packets row-major), lanes [n_avx, n_avx + n_sc) scalar pixels.
struct RtGrid {
    uint32_t w, h, tx, ty, missing, n_avx, n_sc, rows_sc;
    __device__ __forceinline__ bool packet_lane(uint32_t g, uint32_t &x, uint32_t &y) const {
        if (g >= n_avx) return false;
        const uint32_t tile = g >> 3, j = g & 7u;
        x = 4 * (tile % tx) + (j & 3u);
        y = 2 * (tile / tx) + (j >> 2);
        return true;
    }
    __device__ __forceinline__ void scalar_pixel(uint32_t s, uint32_t &x, uint32_t &y) const {
        const uint32_t per = missing + 1;
        if (s < rows_sc * per) {
            y = s / per;
            x = w - 1 - missing + s % per;
        } else {
            y = h - 1;
            x = s - rows_sc * per;
        }
    }
    // Packet pixels later overwritten by the scalar pass (pray.cpp:466-478).
    __device__ __forceinline__ bool overwritten(uint32_t x, uint32_t y) const {
        return y + 1 < h && x + 1 + missing >= w;
    }
};

struct RtArgs {
    RtGrid g;
    const float *rays;      // w*h primary rays
    const int32_t *hit;     // w*h
    const float *t;         // w*h
    const float *lights;    // nl x 6
    uint32_t nl;
    float bg[3];
    float *srays;           // nl x (n_avx + n_sc) shadow rays
    float *slim;            // nl x (n_avx + n_sc) squared limits
    const int32_t *socc;    // nl x (n_avx + n_sc) any-hit results
    uint8_t *rgb;           // w*h*3
};

__device__ __forceinline__ void hit_point(const RtArgs &A, uint32_t p, float &x, float &y, float &z) {
    const float *r = A.rays + 6 * (size_t)p;
    const float t = A.t[p];
    x = r[0] + r[3] * t;
    y = r[1] + r[4] * t;
    z = r[2] + r[5] * t;
}

// One shadow ray per (light, lane): direction and limit of the path the lane
// follows; -1 marks lanes that shoot no shadow ray.
static __global__ void k_rt_shadow_setup(DevScene s, RtArgs A) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lanes = A.g.n_avx + A.g.n_sc;
    if (g >= lanes) return;
    uint32_t x, y;
    const bool avx = A.g.packet_lane(g, x, y);
    if (!avx) A.g.scalar_pixel(g - A.g.n_avx, x, y);
    const uint32_t p = x + y * A.g.w;
    const int32_t h = A.hit[p];
    float px = 0, py = 0, pz = 0, nx = 0, ny = 0, nz = 0;
    if (h >= 0) {
        hit_point(A, p, px, py, pz);
        nx = s.normal[3 * (size_t)h], ny = s.normal[3 * (size_t)h + 1], nz = s.normal[3 * (size_t)h + 2];
    }
    for (uint32_t l = 0; l < A.nl; l++) {
        const size_t at = (size_t)l * lanes + g;
        float *o = A.srays + 6 * at;
        float lim = -1.0f;
        if (h >= 0) {
            float lx = A.lights[6 * l] - px, ly = A.lights[6 * l + 1] - py, lz = A.lights[6 * l + 2] - pz;
            const float d2 = lx * lx + ly * ly + lz * lz;
            if (avx) {
                const float inv = 1.0f / sqrtf(d2);
                lx *= inv, ly *= inv, lz *= inv;
            } else {
                prd::normalize3(lx, ly, lz);
            }
            const float lambert = lx * nx + ly * ny + lz * nz;
            if (lambert >= 0.f) {
                if (!avx) prd::normalize3(lx, ly, lz);  // Ray(intersectionPos, L)
                o[0] = px, o[1] = py, o[2] = pz, o[3] = lx, o[4] = ly, o[5] = lz;
                lim = avx ? d2 : nextafterf(d2, -INFINITY);
            }
        }
        if (lim < 0.f) o[0] = o[1] = o[2] = o[3] = o[4] = o[5] = 0.f, o[5] = 1.f;
        A.slim[at] = lim;
    }
}

static __global__ __launch_bounds__(WAVE) void k_rt_shade(DevScene s, RtArgs A) {
    const uint32_t g = blockIdx.x * WAVE + threadIdx.x;
    const uint32_t lanes = A.g.n_avx + A.g.n_sc;
    const uint32_t lane = threadIdx.x, grp = lane & ~7u;
    const bool valid = g < lanes;
    uint32_t x = 0, y = 0;
    const bool avx = valid && A.g.packet_lane(g, x, y);
    if (valid && !avx) A.g.scalar_pixel(g - A.g.n_avx, x, y);
    const uint32_t p = x + y * A.g.w;
    const int32_t h = valid ? A.hit[p] : -1;
    const bool hit = h >= 0;
    float px = 0, py = 0, pz = 0, nx = 0, ny = 0, nz = 0, ar = 0, ag = 0, ab = 0;
    if (hit) {
        hit_point(A, p, px, py, pz);
        nx = s.normal[3 * (size_t)h], ny = s.normal[3 * (size_t)h + 1], nz = s.normal[3 * (size_t)h + 2];
        ar = s.color[3 * (size_t)h], ag = s.color[3 * (size_t)h + 1], ab = s.color[3 * (size_t)h + 2];
    }
    auto group = [&](bool pred) -> uint32_t { return (uint32_t)((__ballot(pred) >> grp) & 0xFFull); };
    const bool packet_hit = group(avx && hit) != 0;
    float cr = (avx && !hit) ? A.bg[0] : 0.f, cg = (avx && !hit) ? A.bg[1] : 0.f, cb = (avx && !hit) ? A.bg[2] : 0.f;
    bool done = !valid || (avx && !packet_hit) || (!avx && !hit);
    for (uint32_t l = 0; l < A.nl; l++) {
        const float lr = A.lights[6 * l + 3], lg = A.lights[6 * l + 4], lb = A.lights[6 * l + 5];
        float lx = A.lights[6 * l] - px, ly = A.lights[6 * l + 1] - py, lz = A.lights[6 * l + 2] - pz;
        const float d2 = lx * lx + ly * ly + lz * lz;
        if (avx) {
            const float inv = 1.0f / sqrtf(d2);
            lx *= inv, ly *= inv, lz *= inv;
        } else {
            prd::normalize3(lx, ly, lz);
        }
        const float lambert = lx * nx + ly * ny + lz * nz;
        const bool occluded = valid && A.socc[(size_t)l * lanes + g] >= 0;
        if (avx) {
            // packet-wide: skip the light unless some lane faces it; then every
            // hit lane that is not occluded adds its (possibly negative) term
            const bool smask = hit && lambert >= 0.f;
            const bool any_s = group(avx && !done && smask) != 0;
            const bool lit = hit && !(smask && occluded);
            const bool any_lit = group(avx && !done && any_s && lit) != 0;
            if (!done && any_s && any_lit && lit) {
                const float decay = lambert / d2;
                cr = cr + ar * (lr * decay);
                cg = cg + ag * (lg * decay);
                cb = cb + ab * (lb * decay);
            }
            const bool white = cr >= 1.f && cg >= 1.f && cb >= 1.f;
            const uint32_t wmask = group(avx && white);
            if (avx && !done && any_s && any_lit && wmask == 0xFFu) done = true;
        } else if (!done) {
            if (lambert < 0.f || occluded) continue;
            const float inv = 1.0f / d2;
            cr = cr + ((ar * lr) * lambert) * inv;
            cg = cg + ((ag * lg) * lambert) * inv;
            cb = cb + ((ab * lb) * lambert) * inv;
            if (cr >= 1.f && cg >= 1.f && cb >= 1.f) done = true;
        }
    }
    if (!valid) return;
    uint8_t *o = A.rgb + 3 * (size_t)p;
    if (avx) {
        if (A.g.overwritten(x, y)) return;
        if (!packet_hit) {
            o[0] = to_byte(A.bg[0]), o[1] = to_byte(A.bg[1]), o[2] = to_byte(A.bg[2]);
        } else {
            o[0] = to_byte(clamp_std(cr)), o[1] = to_byte(clamp_std(cg)), o[2] = to_byte(clamp_std(cb));
        }
    } else if (!hit) {
        o[0] = to_byte(A.bg[0]), o[1] = to_byte(A.bg[1]), o[2] = to_byte(A.bg[2]);
    } else {
        o[0] = to_byte(clamp_std(cr)), o[1] = to_byte(clamp_std(cg)), o[2] = to_byte(clamp_std(cb));
    }
}

static __global__ void k_camera(Cam c, uint32_t w, uint32_t h, float *rays) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    float dx, dy, dz;
    camera_dir(c, i % w, i / w, dx, dy, dz);
    float *o = rays + 6 * (size_t)i;
    o[0] = c.px, o[1] = c.py, o[2] = c.pz, o[3] = dx, o[4] = dy, o[5] = dz;
}

// ------------------------------------------------------------- path trace --

struct PtArgs {
    uint32_t n_pix, max_depth, rr_depth, k0, k1;
    float bg[3];
    const uint32_t *sample;   // first sample index of the pass (device)
    // Samples in flight together: path p of a pass (a frame of the cache
    // tracer) is sample *sample + p / n_pix of pixel p % n_pix (1 for the
    // batch tracer, whose paths are pixels).
    uint32_t samples = 1;
    // The reference's diffuse estimator (pray.cpp:315-340, chosen at
    // max_depth <= 1, scene_preprocessor.cpp:34-40): the first bounce keeps
    // throughput 1 and no roulette, so the samples' sum is the sum of the
    // radiance arriving at the primary hit, which k_pt_finalize_diffuse
    // multiplies by the hit's albedo once.
    uint32_t diffuse = 0;
};

// Queue entries: ray (6 floats) + state float4 {throughput rgb, path bits}.
// The first bounce's queue holds only states, paths p < n_paths: path p
// starts with the camera ray of pixel p % n_pix, whose hit the renderer
// traces once per call for all samples (k_pt_shade's `prim`), since every
// sample of a pixel starts with the same ray (the reference's primRays,
// pray.cpp:368-372).
static __global__ void k_pt_generate(uint32_t n_paths, float4 *state, uint32_t *count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *count = n_paths;
    if (i >= n_paths) return;
    state[i] = make_float4(1.f, 1.f, 1.f, __uint_as_float(i));
}

// Paths that end in the radiance cache (PT_CACHE) share the plain shading;
// at a non-emissive hit of bounce `cache_bounce`:
//  - while the cache is on, the path stops there and leaves a query
//    {features, weight = throughput * albedo, pixel}: the cache's prediction
//    of the radiance arriving at that point (divided by its albedo)
//    completes the pixel; otherwise the path continues as a plain path;
//  - if the pixel is one of this frame's training pixels (pixels
//    off + k * stride, k < n_train, off drawn per frame), the point's
//    features become training sample k and a training ray leaves it in a
//    direction of its own random stream with throughput 1.  The training
//    ray's hit gives the sample's target (k_pt_train_targets): emission or
//    background, or albedo times the cache's prediction there (the cache
//    trains on its own predictions one bounce further, as in Mueller et al.
//    2021, "Real-time Neural Radiance Caching for Path Tracing").
enum : int { PT_PLAIN = 0, PT_CACHE = 1 };

struct CacheArgs {
    uint32_t cache_bounce;
    const uint32_t *cache_on;   // device flag
    float *queries;             // 16 floats per query
    uint32_t *n_queries;
    uint32_t n_train, stride;   // training pixels per frame, their spacing
    float *feats;               // n_train x 10: training sample features
    float *train_rays;          // training rays (6 floats) ...
    float4 *train_state;        // ... and {1, 1, 1, sample index bits}
    uint32_t *n_train_rays;
    float lo[3], inv[3];        // position -> [0,1]^3
};

__device__ __forceinline__ void sph01(float x, float y, float z, float &th, float &ph) {
    th = acosf(fminf(fmaxf(z, -1.f), 1.f)) * 0.318309886183790672f;
    ph = atan2f(y, x) * 0.159154943091895336f + 0.5f;
}

__device__ __forceinline__ void cache_features(const CacheArgs &C, float px, float py, float pz, float dx, float dy,
                                               float dz, float nx, float ny, float nz, float ar, float ag, float ab,
                                               float *f) {
    f[0] = (px - C.lo[0]) * C.inv[0], f[1] = (py - C.lo[1]) * C.inv[1], f[2] = (pz - C.lo[2]) * C.inv[2];
    sph01(dx, dy, dz, f[3], f[4]);
    sph01(nx, ny, nz, f[5], f[6]);
    f[7] = ar, f[8] = ag, f[9] = ab;
}

// First training pixel of a frame.
__device__ __forceinline__ uint32_t train_offset(uint32_t frame, uint32_t n_pix, uint32_t k0, uint32_t k1) {
    const uint4 u = prd::philox4x32_10(make_uint4(frame, 0u, 0xFFFFFFFFu, 3u), k0, k1);
    return (uint32_t)(((uint64_t)u.x * n_pix) >> 32);
}

// Block-aggregated slot in a queue: one atomic per 256-thread block (every
// thread of the block must call it; `lds` holds 5 words).
__device__ __forceinline__ uint32_t block_slot(bool want, uint32_t *counter, uint32_t *lds) {
    const uint64_t m = __ballot(want);
    const uint32_t lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    if (lane == 0) lds[wv] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (uint32_t k = 0; k < blockDim.x / WAVE; ++k) {
            const uint32_t c = lds[k];
            lds[k] = tot;
            tot += c;
        }
        lds[4] = tot ? atomicAdd(counter, tot) : 0u;
    }
    __syncthreads();
    const uint32_t base = lds[4] + lds[wv] + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    __syncthreads();
    return base;
}

// Wave-aggregated slot in a queue: one atomic per wave.
__device__ __forceinline__ uint32_t wave_slot(bool want, uint32_t *counter) {
    const uint64_t m = __ballot(want);
    uint32_t base = 0;
    if (m) {
        const uint32_t lane = threadIdx.x & (WAVE - 1);
        const uint32_t first = (uint32_t)(__ffsll((unsigned long long)m) - 1);
        if (lane == first) base = atomicAdd(counter, (uint32_t)__popcll(m));
        base = __shfl(base, (int)first);
        base += (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    }
    return base;
}

// Batched rendering of several scenes in one wavefront: queue entry i
// belongs to scene scene_in[i]; pixel indices in the state are global
// (scene's offset + pixel), random numbers use the scene's own key and pixel
// index, so every scene's image equals its single-scene render.
struct BatchScene {
    float bg[3];
    uint32_t k0, k1, pix_off, n_pix, pad;
};

struct BatchArgs {
    const DevScene *table;
    const BatchScene *scenes;
    const uint32_t *scene_in;
    uint32_t *scene_out;
};

// prim: the first bounce, whose ray, hit and distance are the camera ray's,
// read per pixel from rays_in, hit and tt (pixel-indexed: the scene-local
// pixel, or the batch's global one) instead of per queue entry.
template <int MODE, bool MULTI = false>
static __global__ __launch_bounds__(256) void k_pt_shade(DevScene s, PtArgs P, CacheArgs C, uint32_t bounce, int prim,
                                                         const float *rays_in, const float4 *state_in,
                                                         const uint32_t *count_in, const int32_t *hit,
                                                         const float *tt, float *rays_out, float4 *state_out,
                                                         uint32_t *count_out, float *contrib, BatchArgs B) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = *count_in;
    const bool active = i < n;
    bool push = false, query = false, train = false;
    float ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0, tr = 0, tg = 0, tb = 0;
    float tdx = 0, tdy = 0, tdz = 0;
    float feat[10];
    uint32_t pix = 0, tk = 0, scn = 0;
    if (active) {
        const float4 st = state_in[i];
        pix = __float_as_uint(st.w);
        const uint8_t *emit = s.emit;
        const float *color = s.color, *normal = s.normal;
        float bg0 = P.bg[0], bg1 = P.bg[1], bg2 = P.bg[2];
        uint32_t k0 = P.k0, k1 = P.k1, lpix = pix, smp = *P.sample;
        if (!MULTI && P.samples > 1) {
            const uint32_t j = pix / P.n_pix;
            lpix = pix - j * P.n_pix;
            smp += j;
        }
        if (MULTI) {
            scn = B.scene_in[i];
            const DevScene &S = B.table[scn];
            const BatchScene &Q = B.scenes[scn];
            emit = S.emit, color = S.color, normal = S.normal;
            bg0 = Q.bg[0], bg1 = Q.bg[1], bg2 = Q.bg[2];
            k0 = Q.k0, k1 = Q.k1, lpix = pix - Q.pix_off;
        }
        const uint32_t at = prim ? (MULTI ? pix : lpix) : i;  // entry of rays_in, hit, tt
        const int32_t h = hit[at];
        float cr = 0.f, cg = 0.f, cb = 0.f;
        bool term = true;
        if (h < 0) {
            cr = st.x * bg0, cg = st.y * bg1, cb = st.z * bg2;
        } else if (emit[h]) {
            cr = st.x * color[3 * (size_t)h], cg = st.y * color[3 * (size_t)h + 1], cb = st.z * color[3 * (size_t)h + 2];
        } else if (bounce < P.max_depth) {
            const float *r = rays_in + 6 * (size_t)at;
            const float t = tt[at];
            ox = r[0] + r[3] * t, oy = r[1] + r[4] * t, oz = r[2] + r[5] * t;
            const float ar = color[3 * (size_t)h], ag = color[3 * (size_t)h + 1], ab = color[3 * (size_t)h + 2];
            const float nx = normal[3 * (size_t)h], ny = normal[3 * (size_t)h + 1], nz = normal[3 * (size_t)h + 2];
            const bool at_cache = MODE == PT_CACHE && bounce == C.cache_bounce;
            if (at_cache && smp == *P.sample) {  // training pixels: from the frame's first sample
                const uint32_t d = (lpix + P.n_pix - train_offset(smp, P.n_pix, P.k0, P.k1)) % P.n_pix;
                if (d % C.stride == 0u && d / C.stride < C.n_train) {
                    tk = d / C.stride;
                    train = true;
                    cache_features(C, ox, oy, oz, r[3], r[4], r[5], nx, ny, nz, ar, ag, ab, C.feats + 10 * (size_t)tk);
                    const uint4 u = prd::philox4x32_10(make_uint4(lpix, smp, bounce, 1u), P.k0, P.k1);
                    prd::sample_hemisphere(nx, ny, nz, prd::u01(u.x), prd::u01(u.y), tdx, tdy, tdz);
                }
            }
            if (at_cache && *C.cache_on) {
                cache_features(C, ox, oy, oz, r[3], r[4], r[5], nx, ny, nz, ar, ag, ab, feat);
                tr = st.x * ar, tg = st.y * ag, tb = st.z * ab;
                query = true;
                term = false;
            } else {
                const uint4 u = prd::philox4x32_10(make_uint4(lpix, smp, bounce, 0u), k0, k1);
                prd::sample_hemisphere(nx, ny, nz, prd::u01(u.x), prd::u01(u.y), dx, dy, dz);
                if (P.diffuse)
                    tr = st.x, tg = st.y, tb = st.z;
                else
                    tr = st.x * ar, tg = st.y * ag, tb = st.z * ab;
                bool live = true;
                if (!P.diffuse && bounce >= P.rr_depth) {
                    const float q = fmaxf(fmaxf(ar, ag), ab);
                    if (q < 1.f) {
                        if (!(prd::u01(u.z) < q))
                            live = false;
                        else
                            tr = tr / q, tg = tg / q, tb = tb / q;
                    }
                }
                term = !live;
                push = live;
            }
        }
        if (term) {
            float *c = contrib + 3 * (size_t)pix;
            c[0] = cr, c[1] = cg, c[2] = cb;
        }
    }
    __shared__ uint32_t qs[5];
    if (MODE == PT_CACHE) {
        const uint32_t q = block_slot(query, C.n_queries, qs);
        if (query) {
            float *o = C.queries + 16 * (size_t)q;
#pragma unroll
            for (int k = 0; k < 10; ++k) o[k] = feat[k];
            o[10] = tr, o[11] = tg, o[12] = tb, o[13] = __uint_as_float(pix);
        }
        const uint32_t tq = block_slot(train, C.n_train_rays, qs);
        if (train) {
            float *o = C.train_rays + 6 * (size_t)tq;
            o[0] = ox, o[1] = oy, o[2] = oz, o[3] = tdx, o[4] = tdy, o[5] = tdz;
            C.train_state[tq] = make_float4(1.f, 1.f, 1.f, __uint_as_float(tk));
        }
    }
    // one atomic per block for the surviving paths
    const uint32_t slot = block_slot(push, count_out, qs);
    if (push) {
        float *o = rays_out + 6 * (size_t)slot;
        o[0] = ox, o[1] = oy, o[2] = oz, o[3] = dx, o[4] = dy, o[5] = dz;
        state_out[slot] = make_float4(tr, tg, tb, __uint_as_float(pix));
        if (MULTI) B.scene_out[slot] = scn;
    }
}

// First-bounce states of all scenes of a batch, one per pixel of their
// concatenated images: entry g (camera ray g of the concatenated primary
// rays) belongs to the scene whose pixel range holds g.
static __global__ void k_pt_generate_batch(uint32_t n_total, const BatchScene *scenes, uint32_t n_scenes,
                                           float4 *state, uint32_t *scene, uint32_t *count) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g == 0) *count = n_total;
    if (g >= n_total) return;
    uint32_t lo = 0, hi = n_scenes - 1;  // last scene with pix_off <= g
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (scenes[mid].pix_off <= g)
            lo = mid;
        else
            hi = mid - 1;
    }
    state[g] = make_float4(1.f, 1.f, 1.f, __uint_as_float(g));
    scene[g] = lo;
}

// Targets of the training samples from their training rays' hits: emission
// or background directly; at a non-emissive hit a query {features, weight =
// albedo, sample index} whose prediction the cache's inference writes.
static __global__ __launch_bounds__(256) void k_pt_train_targets(DevScene s, PtArgs P, CacheArgs C,
                                                                 const float *rays, const float4 *state,
                                                                 const uint32_t *count, const int32_t *hit,
                                                                 const float *tt, float4 *targets, float *queries,
                                                                 uint32_t *n_queries) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i < *count;
    bool query = false;
    float feat[10], wr = 0.f, wg = 0.f, wb = 0.f;
    uint32_t k = 0;
    if (active) {
        k = __float_as_uint(state[i].w);
        const int32_t h = hit[i];
        if (h < 0) {
            targets[k] = make_float4(P.bg[0], P.bg[1], P.bg[2], 1.f);
        } else if (s.emit[h]) {
            targets[k] = make_float4(s.color[3 * (size_t)h], s.color[3 * (size_t)h + 1], s.color[3 * (size_t)h + 2], 1.f);
        } else {
            const float *r = rays + 6 * (size_t)i;
            const float t = tt[i];
            wr = s.color[3 * (size_t)h], wg = s.color[3 * (size_t)h + 1], wb = s.color[3 * (size_t)h + 2];
            cache_features(C, r[0] + r[3] * t, r[1] + r[4] * t, r[2] + r[5] * t, r[3], r[4], r[5],
                           s.normal[3 * (size_t)h], s.normal[3 * (size_t)h + 1], s.normal[3 * (size_t)h + 2], wr, wg,
                           wb, feat);
            query = true;
        }
    }
    const uint32_t q = wave_slot(query, n_queries);
    if (query) {
        float *o = queries + 16 * (size_t)q;
#pragma unroll
        for (int j = 0; j < 10; ++j) o[j] = feat[j];
        o[10] = wr, o[11] = wg, o[12] = wb, o[13] = __uint_as_float(k);
    }
}

// sum += contribution of this pass (sample order), then advance the sample.
// Adds the pass's `samples` contributions per pixel to the running sum, in
// sample order (the order the one-sample passes would add them in), and
// advances the sample index past them.
static __global__ void k_pt_accumulate(uint32_t n_pix, uint32_t samples, const float *contrib, float *sum,
                                       uint32_t *sample) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 3 * n_pix) {
        float v = sum[i];
        for (uint32_t j = 0; j < samples; ++j) v = v + contrib[(size_t)j * 3 * n_pix + i];
        sum[i] = v;
    }
    if (i == 0) *sample = *sample + samples;
}

// Adds a queue count to a traced-ray counter (inside graphs: a kernel, not
// a memset or host call).
static __global__ void k_acc_rays(unsigned long long *acc, const uint32_t *count) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *acc += *count;
}

static __global__ void k_pt_finalize(uint32_t n_pix, const float *sum, float spp, float *hdr, uint8_t *rgb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n_pix) return;
    const float m = sum[i] / spp;
    if (hdr) hdr[i] = m;
    if (rgb) rgb[i] = to_byte(clamp_std(m));
}

// The diffuse estimator's pixels (pray.cpp:315-340 with the depth limit of
// at most 1 that selects it): the camera ray's miss gives the background and
// an emissive hit its emission, both once, not averaged; a non-emissive hit
// gives black at depth limit 0 and otherwise (sum of the samples' arriving
// radiance * albedo) / samples, multiplied and divided in that order.
// Pixels [0, n_pix) of one scene; hit is that scene's camera-ray hits.
static __global__ void k_pt_finalize_diffuse(uint32_t n_pix, const float *sum, float spp, const int32_t *hit,
                                             const uint8_t *emit, const float *color, float bg0, float bg1, float bg2,
                                             uint32_t max_depth, float *hdr, uint8_t *rgb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n_pix) return;
    const uint32_t c = i % 3;
    const int32_t h = hit[i / 3];
    float m;
    if (h < 0)
        m = c == 0 ? bg0 : (c == 1 ? bg1 : bg2);
    else if (emit[h])
        m = color[3 * (size_t)h + c];
    else if (max_depth == 0)
        m = 0.f;
    else
        m = (sum[i] * color[3 * (size_t)h + c]) / spp;
    if (hdr) hdr[i] = m;
    if (rgb) rgb[i] = to_byte(clamp_std(m));
}

static __global__ void k_fill_f32(float *p, uint32_t n, float v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace prr
