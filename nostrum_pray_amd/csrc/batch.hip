// Batched multi-scene path tracing and the traversal autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "../../include/pray_hip.h"
#include "context.hpp"
#include "render.hpp"

namespace prc {
namespace {

// Rays for timing: origins uniform in the scene box grown by 10%, directions
// uniform on the sphere (the incoherent stream of pray_gen_rays).
int sample_rays(pray_ctx *c, const SceneRec &sc, uint32_t n, uint64_t seed, float *d_rays) {
    float box6[6];
    for (int k = 0; k < 3; k++) {
        const float ext = sc.d.box[k + 3] - sc.d.box[k];
        box6[k] = sc.d.box[k] - 0.05f * ext;
        box6[k + 3] = 1.1f * ext;
    }
    return pray_gen_rays_device(c, seed, box6, n, d_rays, c->stream);
}

float time_trace(pray_ctx *c, SceneRec &sc, const Tuning &tu, const float *rays, uint32_t n, float eps,
                 int32_t *idx, float *t, int reps) {
    const Tuning saved = sc.tune;
    sc.tune = tu;
    float best = 1e30f;
    hipEvent_t e0 = c->ievents[0], e1 = c->ievents[1];
    for (int r = 0; r < reps + 1; r++) {
        if (hipEventRecord(e0, c->stream) != hipSuccess) break;
        if (enqueue_trace(c, sc, rays, n, nullptr, PRAY_TRACE_CLOSEST, eps, nullptr, idx, t, nullptr, c->stream))
            break;
        if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) break;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) break;
        if (r > 0) best = std::min(best, ms);  // the first run warms caches and code
    }
    sc.tune = saved;
    return best;
}

}  // namespace
}  // namespace prc

using namespace prc;

extern "C" {

int pray_autotune(pray_ctx *c, int scene_id, uint32_t n_rays, int force, int *tuning) {
    if (!c) return PRAY_ERR_INVALID;
    SceneRec *sc = find_scene(c, scene_id);
    if (!sc) return fail(c, PRAY_ERR_NOSCENE, "unknown scene %d", scene_id);
    if (sc->d.n_tris == 0) {
        if (tuning) tuning_values(sc->tune, tuning);
        return PRAY_OK;
    }
    HIPC(c, hipSetDevice(c->device));
    const int bucket = size_bucket(sc->d.n_tris);
    auto it = c->tuned.find(bucket);
    if (!force && it != c->tuned.end()) {
        sc->tune = it->second;
        // the bucket's node form where this tree has it, else its own
        const int nl = sc->tune.node_levels;
        if ((nl == 2 && !sc->d.quads) || (nl == 3 && !sc->d.boxes) || (nl == 4 && !sc->d.boxes4))
            sc->tune.node_levels = sc->d.boxes4 ? 4 : (sc->d.boxes ? 3 : 1);
    } else {
        if (n_rays == 0) n_rays = 1u << 20;
        for (int k = 0; k < 2; k++)
            if (!c->ievents[k]) HIPC(c, hipEventCreate(&c->ievents[k]));
        DBuf rays, idx, tt;
        HIPC(c, hipMalloc(&rays.p, (size_t)n_rays * 24));
        HIPC(c, hipMalloc(&idx.p, (size_t)n_rays * 4));
        HIPC(c, hipMalloc(&tt.p, (size_t)n_rays * 4));
        int rc = sample_rays(c, *sc, n_rays, 0x5EEDull + (uint64_t)scene_id, (float *)rays.p);
        if (rc) return rc;
        // Grid over node records x stack depth x leaf ILP x refill
        // threshold at full occupancy; then leaf batching (with two refill
        // thresholds), the workgroup size and the occupancy of the winner
        // (packed layout).
        Tuning best;
        float best_ms = 1e30f;
        auto consider = [&](Tuning cand) {
            const float ms = time_trace(c, *sc, cand, (const float *)rays.p, n_rays, 1e-9f, (int32_t *)idx.p,
                                        (float *)tt.p, 2);
            if (ms < best_ms) best_ms = ms, best = cand;
        };
        for (int nl : {1, 2, 3, 4}) {
            if ((nl == 2 && !sc->d.quads) || (nl == 3 && !sc->d.boxes) || (nl == 4 && !sc->d.boxes4)) continue;
            for (int K : {4, 6, 8, 16})
                for (int ilp : {1, 2})
                    for (int fm : {16, 24, 32, 48}) {
                        Tuning t;
                        t.stack_lds = K, t.leaf_ilp = ilp, t.fetch_min = fm, t.node_levels = nl;
                        consider(t);
                    }
        }
        for (int ilp : {1, 2}) {
            const Tuning base = best;
            for (int lb : {16, 24, 32})
                for (int fm : {16, 24}) {
                    Tuning t = base;
                    t.leaf_ilp = ilp, t.leaf_batch = lb, t.fetch_min = fm;
                    consider(t);
                }
        }
        const Tuning grid_best = best;
        for (int wg : {2, 4}) {
            Tuning t = grid_best;
            t.wg_waves = wg;
            consider(t);
        }
        for (int occ : {75, 50}) {
            Tuning t = best;
            t.occupancy = occ;
            consider(t);
        }

        HIPC(c, hipGetLastError());
        c->tuned[bucket] = best;
        sc->tune = best;
    }
    if (tuning) {
        tuning_values(sc->tune, tuning);
    }
    return PRAY_OK;
}

int pray_autotune_pt(pray_ctx *c, int scene_id, const float *cam16, uint32_t w, uint32_t h, uint32_t max_depth,
                     uint32_t rr_depth, uint64_t seed, const float *bg, float det_eps, int *tuning) {
    if (!c) return PRAY_ERR_INVALID;
    SceneRec *sc = find_scene(c, scene_id);
    if (!sc) return fail(c, PRAY_ERR_NOSCENE, "unknown scene %d", scene_id);
    if (!cam16 || !bg) return fail(c, PRAY_ERR_INVALID, "null argument");
    if ((uint64_t)w * h == 0) return fail(c, PRAY_ERR_INVALID, "empty image");
    if (sc->d.n_tris) {
        // Same grid as pray_autotune, timed on two-sample renders of the
        // caller's image (host clock around the whole call, which allocates
        // its buffers; the fastest of two runs after one warm-up).
        const Tuning saved = sc->tune;
        Tuning best = saved;
        double best_ms = 1e300;
        // The grid takes its fastest point; the later stages (leaf batching,
        // workgroup size, occupancy) change the launch shape of a wavefront
        // that in a batch is far larger than this scene's two samples, so
        // they replace the grid's point only when at least 3% faster here.
        double margin = 1.0;
        auto consider = [&](Tuning cand) -> int {
            sc->tune = cand;
            double fastest = 1e300;
            for (int r = 0; r < 3; r++) {
                const auto t0 = std::chrono::steady_clock::now();
                const int rc = pray_render_pt_ex(c, scene_id, cam16, w, h, 2, max_depth, rr_depth, seed, bg, det_eps,
                                                 nullptr, nullptr, 1);
                if (rc) return rc;
                const double ms =
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                if (r > 0) fastest = std::min(fastest, ms);
            }
            if (fastest < best_ms * margin) best_ms = fastest, best = cand;
            return PRAY_OK;
        };
        int rc = PRAY_OK;
        for (int nl : {1, 2, 3, 4}) {
            if ((nl == 2 && !sc->d.quads) || (nl == 3 && !sc->d.boxes) || (nl == 4 && !sc->d.boxes4)) continue;
            for (int K : {4, 6, 8, 16})
                for (int ilp : {1, 2})
                    for (int fm : {16, 24, 32, 48}) {
                        Tuning t;
                        t.stack_lds = K, t.leaf_ilp = ilp, t.fetch_min = fm, t.node_levels = nl;
                        if (!rc) rc = consider(t);
                    }
        }
        margin = 0.97;
        for (int ilp : {1, 2}) {
            const Tuning base = best;
            for (int lb : {16, 24, 32})
                for (int fm : {16, 24}) {
                    Tuning t = base;
                    t.leaf_ilp = ilp, t.leaf_batch = lb, t.fetch_min = fm;
                    if (!rc) rc = consider(t);
                }
        }
        const Tuning grid_best = best;
        for (int wg : {2, 4}) {
            Tuning t = grid_best;
            t.wg_waves = wg;
            if (!rc) rc = consider(t);
        }
        for (int occ : {75, 50}) {
            Tuning t = best;
            t.occupancy = occ;
            if (!rc) rc = consider(t);
        }
        if (rc) {
            sc->tune = saved;
            return rc;
        }
        sc->tune = best;
    }
    if (tuning) tuning_values(sc->tune, tuning);
    return PRAY_OK;
}

int pray_render_pt_batch(pray_ctx *c, uint32_t n_scenes, const int *scene_ids, const float *cams16,
                         const uint32_t *widths, const uint32_t *heights, uint32_t spp, uint32_t max_depth,
                         uint32_t rr_depth, const uint64_t *seeds, const float *backgrounds, float det_eps,
                         float *const *hdr, uint8_t *const *rgb, float *gpu_ms) {
    if (!c) return PRAY_ERR_INVALID;
    if (n_scenes == 0 || spp == 0) return PRAY_OK;
    if (!scene_ids || !cams16 || !widths || !heights || !seeds || !backgrounds)
        return fail(c, PRAY_ERR_INVALID, "null argument");
    // As pray_render_pt: bounded bounce loop; the mean divides by the sample
    // count as a float, which must stay exact.
    if (max_depth > 64) return fail(c, PRAY_ERR_INVALID, "max_depth must be at most 64");
    if (spp > (1u << 24)) return fail(c, PRAY_ERR_INVALID, "more than 2^24 samples");
    HIPC(c, hipSetDevice(c->device));
    std::vector<prt::DevScene> table(n_scenes);
    std::vector<prr::BatchScene> bs(n_scenes);
    uint64_t total = 0;
    uint32_t tree_depth = 1;
    std::vector<uint64_t> bucket_pixels(16, 0);
    std::vector<uint64_t> bucket_best(16, 0);          // pixels of each bucket's largest image
    std::vector<const SceneRec *> bucket_scene(16, nullptr);  // and its scene
    for (uint32_t k = 0; k < n_scenes; k++) {
        SceneRec *sc = find_scene(c, scene_ids[k]);
        if (!sc) return fail(c, PRAY_ERR_NOSCENE, "unknown scene %d", scene_ids[k]);
        if (sc->d.n_tris == 0 || !sc->d.recs) return fail(c, PRAY_ERR_INVALID, "empty scene %d in a batch", k);
        if (!sc->d.normal || !sc->d.color || !sc->d.emit)
            return fail(c, PRAY_ERR_INVALID, "scene %d was uploaded without shading data", k);
        const uint64_t np = (uint64_t)widths[k] * heights[k];
        if (np == 0) return fail(c, PRAY_ERR_INVALID, "scene %d has an empty image", k);
        table[k] = sc->d;
        prr::BatchScene &b = bs[k];
        b.bg[0] = backgrounds[3 * k], b.bg[1] = backgrounds[3 * k + 1], b.bg[2] = backgrounds[3 * k + 2];
        b.k0 = (uint32_t)seeds[k], b.k1 = (uint32_t)(seeds[k] >> 32);
        b.pix_off = (uint32_t)total;
        b.n_pix = (uint32_t)np;
        b.pad = 0;
        total += np;
        if (total > (1u << 28)) return fail(c, PRAY_ERR_INVALID, "batch too large (more than 2^28 pixels)");
        tree_depth = std::max(tree_depth, sc->d.max_depth);
        const int sb = size_bucket(sc->d.n_tris);
        bucket_pixels[sb] += np;
        if (np > bucket_best[sb]) bucket_best[sb] = np, bucket_scene[sb] = sc;
    }
    // One launch per bounce serves every scene: it runs with the setting of
    // the scene with the largest image in the size bucket that holds most of
    // the pixels (its autotuned bucket setting after pray_autotune, or its
    // own after pray_autotune_pt / pray_set_tuning).
    const int main_bucket = (int)(std::max_element(bucket_pixels.begin(), bucket_pixels.end()) - bucket_pixels.begin());
    Tuning tu = bucket_scene[main_bucket]->tune;
    tu.layout = PRAY_LAYOUT_PACKED;

    // Buffers from the context's render workspace (slots 40..58), reused by
    // later calls.
    struct Buf {
        void *p = nullptr;
    } dtable, dbs, prim, rq[2], sq[2], qs[2], cnt, hit, tt, contrib, sum, outh, outb, rays_acc, phit, pt;
    WS_GET(c, 40, dtable.p, n_scenes * sizeof(prt::DevScene));
    WS_GET(c, 41, dbs.p, n_scenes * sizeof(prr::BatchScene));
    WS_GET(c, 42, prim.p, total * 24);
    for (int q = 0; q < 2; q++) {
        WS_GET(c, 43 + q, rq[q].p, total * 24);
        WS_GET(c, 45 + q, sq[q].p, total * 16);
        WS_GET(c, 47 + q, qs[q].p, total * 4);
    }
    WS_GET(c, 49, cnt.p, 4 * sizeof(uint32_t));
    WS_GET(c, 50, hit.p, total * 4);
    WS_GET(c, 51, tt.p, total * 4);
    WS_GET(c, 52, contrib.p, total * 12);
    WS_GET(c, 53, sum.p, total * 12);
    WS_GET(c, 54, outh.p, total * 12);
    WS_GET(c, 55, rays_acc.p, 8);
    if (rgb) WS_GET(c, 56, outb.p, total * 3);
    WS_GET(c, 57, phit.p, total * 4);  // the camera rays' hits, shared by all samples
    WS_GET(c, 58, pt.p, total * 4);
    hipStream_t st = c->stream;
    HIPC(c, hipMemcpyAsync(dtable.p, table.data(), n_scenes * sizeof(prt::DevScene), hipMemcpyHostToDevice, st));
    HIPC(c, hipMemcpyAsync(dbs.p, bs.data(), n_scenes * sizeof(prr::BatchScene), hipMemcpyHostToDevice, st));
    for (uint32_t k = 0; k < n_scenes; k++) {
        prr::Cam cam;
        std::memcpy(&cam, cams16 + 16 * (size_t)k, sizeof cam);
        const uint64_t np = (uint64_t)widths[k] * heights[k];
        hipLaunchKernelGGL(prr::k_camera, dim3(nblocks(np, 256)), dim3(256), 0, st, cam, widths[k], heights[k],
                           (float *)prim.p + 6 * (size_t)bs[k].pix_off);
    }
    hipLaunchKernelGGL(prr::k_fill_f32, dim3(nblocks(3 * total, 256)), dim3(256), 0, st, (float *)sum.p,
                       (uint32_t)(3 * total), 0.f);
    hipLaunchKernelGGL(prt::k_zero_u32, dim3(1), dim3(64), 0, st, (uint32_t *)cnt.p, 4u);
    hipLaunchKernelGGL(prt::k_zero_u32, dim3(1), dim3(64), 0, st, (uint32_t *)rays_acc.p, 2u);
    HIPC(c, hipGetLastError());
    uint32_t *count = (uint32_t *)cnt.p;
    prr::PtArgs P{};
    P.n_pix = (uint32_t)total;
    P.max_depth = max_depth;
    P.rr_depth = rr_depth;
    P.sample = count + 2;
    P.diffuse = max_depth <= 1 ? 1u : 0u;  // the reference's estimator choice (scene_preprocessor.cpp:34-40)
    const prt::DevScene *dtab = (const prt::DevScene *)dtable.p;
    int rc = prepare_trace_multi(c, tu, tree_depth, total);
    if (rc) return rc;
    // Every sample of a pixel starts with the pixel's camera ray: its hit is
    // traced once for the whole call (and counted once among the rays), with
    // the pixels' scenes from a first generate.
    hipLaunchKernelGGL(prr::k_pt_generate_batch, dim3(nblocks(total, 256)), dim3(256), 0, st, (uint32_t)total,
                       (const prr::BatchScene *)dbs.p, n_scenes, (float4 *)sq[0].p, (uint32_t *)qs[0].p, count + 0);
    hipLaunchKernelGGL(prr::k_acc_rays, dim3(1), dim3(64), 0, st, (unsigned long long *)rays_acc.p,
                       (const uint32_t *)(count + 0));
    rc = enqueue_trace_multi(c, dtab, (const uint32_t *)qs[0].p, tree_depth, tu, (const float *)prim.p, total,
                             nullptr, det_eps, (int32_t *)phit.p, (float *)pt.p, st);
    if (rc) return rc;

    auto pass = [&]() -> int {
        Range rp("batch pass");
        hipLaunchKernelGGL(prr::k_pt_generate_batch, dim3(nblocks(total, 256)), dim3(256), 0, st, (uint32_t)total,
                           (const prr::BatchScene *)dbs.p, n_scenes, (float4 *)sq[0].p, (uint32_t *)qs[0].p,
                           count + 0);
        for (uint32_t b = 0; b <= max_depth; b++) {
            const int cur = b & 1, nxt = cur ^ 1;
            const bool first_bounce = b == 0;  // hits of the camera rays, traced once
            hipLaunchKernelGGL(prt::k_zero_u32, dim3(1), dim3(64), 0, st, count + nxt, 1u);
            if (!first_bounce) {
                hipLaunchKernelGGL(prr::k_acc_rays, dim3(1), dim3(64), 0, st, (unsigned long long *)rays_acc.p,
                                   (const uint32_t *)(count + cur));
                Range r("batch extend");
                const int e = enqueue_trace_multi(c, dtab, (const uint32_t *)qs[cur].p, tree_depth, tu,
                                                  (const float *)rq[cur].p, total, count + cur, det_eps,
                                                  (int32_t *)hit.p, (float *)tt.p, st);
                if (e) return e;
            }
            Range rs("batch shade");
            prr::BatchArgs B{dtab, (const prr::BatchScene *)dbs.p, (const uint32_t *)qs[cur].p, (uint32_t *)qs[nxt].p};
            hipLaunchKernelGGL((prr::k_pt_shade<prr::PT_PLAIN, true>), dim3(nblocks(total, 256)), dim3(256), 0, st,
                               prt::DevScene{}, P, prr::CacheArgs{}, b, first_bounce ? 1 : 0,
                               (const float *)(first_bounce ? prim.p : rq[cur].p), (const float4 *)sq[cur].p,
                               (const uint32_t *)(count + cur), (const int32_t *)(first_bounce ? phit.p : hit.p),
                               (const float *)(first_bounce ? pt.p : tt.p), (float *)rq[nxt].p, (float4 *)sq[nxt].p,
                               count + nxt, (float *)contrib.p, B);
        }
        hipLaunchKernelGGL(prr::k_pt_accumulate, dim3(nblocks(3 * total, 256)), dim3(256), 0, st, (uint32_t)total, 1u,
                           (const float *)contrib.p, (float *)sum.p, count + 2);
        HIPC(c, hipGetLastError());
        return PRAY_OK;
    };
    for (int k = 2; k < 4; k++)
        if (!c->ievents[k]) HIPC(c, hipEventCreate(&c->ievents[k]));
    if (c->use_graphs) {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        HIPC(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        rc = pass();
        hipError_t e = hipStreamEndCapture(st, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return fail(c, PRAY_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return fail(c, PRAY_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
        }
        // The work enqueued before the capture (camera rays and their hits)
        // completes before the first replay: the replays are not left to
        // order themselves after it.
        e = hipStreamSynchronize(st);
        if (e == hipSuccess) e = hipEventRecord(c->ievents[2], st);
        for (uint32_t s = 0; s < spp && e == hipSuccess; s++) e = hipGraphLaunch(exec, st);
        if (e == hipSuccess) e = hipEventRecord(c->ievents[3], st);
        const hipError_t e2 = hipStreamSynchronize(st);
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess || e2 != hipSuccess)
            return fail(c, PRAY_ERR_HIP, "graph launch failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    } else {
        // the same passes enqueued kernel by kernel
        HIPC(c, hipEventRecord(c->ievents[2], st));
        for (uint32_t s = 0; s < spp; s++) {
            rc = pass();
            if (rc) return rc;
        }
        HIPC(c, hipEventRecord(c->ievents[3], st));
        HIPC(c, hipStreamSynchronize(st));
    }
    if (gpu_ms) HIPC(c, hipEventElapsedTime(gpu_ms, c->ievents[2], c->ievents[3]));
    HIPC(c, hipMemcpy(&c->last_rays, rays_acc.p, 8, hipMemcpyDeviceToHost));
    c->last_queues.clear();
    if (P.diffuse) {
        for (uint32_t k = 0; k < n_scenes; k++) {
            const size_t off = bs[k].pix_off, np = bs[k].n_pix;
            hipLaunchKernelGGL(prr::k_pt_finalize_diffuse, dim3(nblocks(3 * np, 256)), dim3(256), 0, st,
                               (uint32_t)np, (const float *)sum.p + 3 * off, (float)spp,
                               (const int32_t *)phit.p + off, table[k].emit, table[k].color, bs[k].bg[0],
                               bs[k].bg[1], bs[k].bg[2], max_depth, (float *)outh.p + 3 * off,
                               outb.p ? (uint8_t *)outb.p + 3 * off : nullptr);
        }
    } else {
        hipLaunchKernelGGL(prr::k_pt_finalize, dim3(nblocks(3 * total, 256)), dim3(256), 0, st, (uint32_t)total,
                           (const float *)sum.p, (float)spp, (float *)outh.p, (uint8_t *)outb.p);
    }
    HIPC(c, hipGetLastError());
    for (uint32_t k = 0; k < n_scenes; k++) {
        const size_t off = bs[k].pix_off, np = bs[k].n_pix;
        if (hdr && hdr[k])
            HIPC(c, hipMemcpyAsync(hdr[k], (float *)outh.p + 3 * off, np * 12, hipMemcpyDeviceToHost, st));
        if (rgb && rgb[k])
            HIPC(c, hipMemcpyAsync(rgb[k], (uint8_t *)outb.p + 3 * off, np * 3, hipMemcpyDeviceToHost, st));
    }
    HIPC(c, hipStreamSynchronize(st));
    return PRAY_OK;
}

int pray_last_render_rays(pray_ctx *c, uint64_t *rays) {
    if (!c || !rays) return PRAY_ERR_INVALID;
    *rays = c->last_rays;
    return PRAY_OK;
}

int pray_last_render_queues(pray_ctx *c, uint64_t *counts, uint32_t n, uint32_t *n_bounces) {
    if (!c) return PRAY_ERR_INVALID;
    if (n_bounces) *n_bounces = (uint32_t)c->last_queues.size();
    for (uint32_t b = 0; counts && b < n && b < c->last_queues.size(); b++) counts[b] = c->last_queues[b];
    return PRAY_OK;
}

}  // extern "C"
