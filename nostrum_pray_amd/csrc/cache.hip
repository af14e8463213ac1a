// Radiance cache: C ABI of the MLP (create, parameters, inference, gradients,
// training steps) and the path tracer whose paths end in the cache.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/pray_hip.h"
#include "context.hpp"
#include "mlp.hpp"
#include "render.hpp"

namespace prc {

namespace {

constexpr uint32_t CHUNK = 1024;  // samples per weight-gradient partial

uint32_t pad64(uint32_t n) { return (n + 63u) & ~63u; }

// Deterministic parameter initialisation (SplitMix64 -> uniform).
struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

struct CacheRec {
    uint32_t max_batch = 0, n_pad_max = 0, max_chunks = 0;
    float lr = 1e-2f, beta1 = 0.9f, beta2 = 0.99f, eps = 1e-8f;
    float *master = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *ffrag = nullptr, *bfrag = nullptr;
    float *feats = nullptr;
    float4 *targets = nullptr;
    uint16_t *xin = nullptr, *act = nullptr, *dz = nullptr;
    float *partial = nullptr, *loss_part = nullptr, *grad = nullptr;
    uint32_t *ctr = nullptr;  // [0] valid samples, [1] steps, [2] cache on
    float *lossv = nullptr;   // [0] last loss, [1] running loss
    uint32_t steps = 0;       // host mirror of ctr[1]
    std::vector<void *> allocs;
};

void free_cache_rec(CacheRec *r) {
    if (!r) return;
    for (void *p : r->allocs) (void)hipFree(p);
    delete r;
}

namespace {

template <typename T>
int alloc(pray_ctx *c, CacheRec &r, T **p, size_t bytes) {
    HIPC(c, hipMalloc((void **)p, bytes));
    r.allocs.push_back((void *)*p);
    return PRAY_OK;
}

CacheRec *find_cache(pray_ctx *c, int id) {
    auto it = c->caches.find(id);
    return it == c->caches.end() ? nullptr : it->second;
}

// Resets the Adam state and step counter and packs the master weights into
// the bf16 fragment arrays.
int reset_optimiser(pray_ctx *c, CacheRec &r) {
    hipStream_t st = c->stream;
    HIPC(c, hipMemsetAsync(r.m, 0, prm::N_PARAMS * 4, st));
    HIPC(c, hipMemsetAsync(r.v, 0, prm::N_PARAMS * 4, st));
    HIPC(c, hipMemsetAsync(r.ctr, 0, 8 * 4, st));
    HIPC(c, hipMemsetAsync(r.lossv, 0, 2 * 4, st));
    hipLaunchKernelGGL(prm::k_mlp_pack, dim3(nblocks(prm::N_PARAMS, 256)), dim3(256), 0, st, (const float *)r.master,
                       r.ffrag, r.bfrag);
    HIPC(c, hipGetLastError());
    r.steps = 0;
    return PRAY_OK;
}

// Forward/backward over the batch in r.feats / r.targets (n samples, r.ctr[0]
// valid), weight-gradient partials; with `update`, an Adam step and the
// step bookkeeping.  Graph-capturable (no host synchronisation, no memsets).
int enqueue_train(pray_ctx *c, CacheRec &r, uint32_t n, bool update, float threshold, uint32_t min_steps,
                  float *hist, uint32_t hist_base, uint32_t hist_cap, hipStream_t st) {
    const uint32_t np = pad64(n);
    if (np > r.n_pad_max) return fail(c, PRAY_ERR_INVALID, "batch of %u exceeds the cache's %u", n, r.max_batch);
    prm::TrainArgs ta;
    ta.feats = r.feats;
    ta.targets = r.targets;
    ta.n = n;
    ta.n_pad = np;
    ta.count = r.ctr;
    ta.xin = r.xin;
    ta.act = r.act;
    ta.dz = r.dz;
    ta.loss_part = r.loss_part;
    hipLaunchKernelGGL(prm::k_mlp_train, dim3(np / 64), dim3(256), 0, st, (const prm::bf16x8 *)r.ffrag,
                       (const prm::bf16x8 *)r.bfrag, ta);
    const uint32_t chunks = (np + CHUNK - 1) / CHUNK;
    prm::WgradArgs wa;
    wa.n_pad = np;
    wa.chunk = CHUNK;
    wa.xin = r.xin;
    wa.act = r.act;
    wa.dz = r.dz;
    wa.partial = r.partial;
    hipLaunchKernelGGL(prm::k_mlp_wgrad, dim3(prm::N_WTILES, chunks), dim3(64), 0, st, wa);
    if (update) {
        prm::AdamArgs aa;
        aa.master = r.master, aa.m = r.m, aa.v = r.v;
        aa.partial = r.partial;
        aa.n_chunks = chunks;
        aa.step = r.ctr + 1;
        aa.lr = r.lr, aa.beta1 = r.beta1, aa.beta2 = r.beta2, aa.eps = r.eps;
        aa.ffrag = r.ffrag, aa.bfrag = r.bfrag;
        hipLaunchKernelGGL(prm::k_mlp_adam, dim3(nblocks(prm::N_PARAMS, 256)), dim3(256), 0, st, aa);
        prm::FinishArgs fa;
        fa.loss_part = r.loss_part;
        fa.n_part = np / 16;
        fa.count = r.ctr;
        fa.step = r.ctr + 1;
        fa.loss_last = r.lossv;
        fa.loss_ema = r.lossv + 1;
        fa.cache_on = r.ctr + 2;
        fa.threshold = threshold;
        fa.min_steps = min_steps;
        fa.hist = hist;
        fa.hist_base = hist_base;
        fa.hist_cap = hist_cap;
        hipLaunchKernelGGL(prm::k_mlp_finish, dim3(1), dim3(256), 0, st, fa);
    } else {
        hipLaunchKernelGGL(prm::k_mlp_reduce_grad, dim3(nblocks(prm::N_PARAMS, 256)), dim3(256), 0, st,
                           (const float *)r.partial, chunks, r.grad);
    }
    HIPC(c, hipGetLastError());
    return PRAY_OK;
}

// Copies a host batch into the cache's training buffers (all samples valid).
int load_batch(pray_ctx *c, CacheRec &r, const float *feats, const float *targets, uint64_t n) {
    if (n == 0 || n > r.max_batch) return fail(c, PRAY_ERR_INVALID, "batch size %llu not in [1, %u]",
                                               (unsigned long long)n, r.max_batch);
    std::vector<float4> t(n);
    for (uint64_t i = 0; i < n; i++) t[i] = make_float4(targets[3 * i], targets[3 * i + 1], targets[3 * i + 2], 1.f);
    const uint32_t cnt = (uint32_t)n;
    hipStream_t st = c->stream;
    HIPC(c, hipMemcpyAsync(r.feats, feats, n * prm::N_FEAT * 4, hipMemcpyHostToDevice, st));
    HIPC(c, hipMemcpyAsync(r.targets, t.data(), n * 16, hipMemcpyHostToDevice, st));
    HIPC(c, hipMemcpyAsync(r.ctr, &cnt, 4, hipMemcpyHostToDevice, st));
    HIPC(c, hipStreamSynchronize(st));
    return PRAY_OK;
}

int batch_loss(pray_ctx *c, CacheRec &r, uint64_t n, float *loss) {
    if (!loss) return PRAY_OK;
    const uint32_t parts = pad64((uint32_t)n) / 16;
    std::vector<float> lp(parts);
    HIPC(c, hipMemcpy(lp.data(), r.loss_part, parts * 4, hipMemcpyDeviceToHost));
    double s = 0;
    for (float x : lp) s += x;
    *loss = (float)(s / (3.0 * (double)n));
    return PRAY_OK;
}

}  // namespace
}  // namespace prc

using namespace prc;

extern "C" {

int pray_cache_create(pray_ctx *c, uint64_t seed, float learning_rate, uint32_t max_batch, int *cache_id) {
    if (!c || !cache_id) return fail(c, PRAY_ERR_INVALID, "null argument");
    if (max_batch == 0 || max_batch > (1u << 22)) return fail(c, PRAY_ERR_INVALID, "max_batch must be in [1, 2^22]");
    if (!(learning_rate > 0.f)) return fail(c, PRAY_ERR_INVALID, "learning rate must be positive");
    HIPC(c, hipSetDevice(c->device));
    auto *r = new CacheRec();
    r->max_batch = max_batch;
    r->n_pad_max = pad64(max_batch);
    r->max_chunks = (r->n_pad_max + CHUNK - 1) / CHUNK;
    r->lr = learning_rate;
    const size_t np = r->n_pad_max;
    int rc = PRAY_OK;
    if (!rc) rc = alloc(c, *r, &r->master, prm::N_PARAMS * 4);
    if (!rc) rc = alloc(c, *r, &r->m, prm::N_PARAMS * 4);
    if (!rc) rc = alloc(c, *r, &r->v, prm::N_PARAMS * 4);
    if (!rc) rc = alloc(c, *r, &r->grad, prm::N_PARAMS * 4);
    if (!rc) rc = alloc(c, *r, &r->ffrag, prm::N_FFRAG * 64 * 16);
    if (!rc) rc = alloc(c, *r, &r->bfrag, prm::N_BFRAG * 64 * 16);
    if (!rc) rc = alloc(c, *r, &r->feats, np * prm::N_FEAT * 4);
    if (!rc) rc = alloc(c, *r, &r->targets, np * 16);
    if (!rc) rc = alloc(c, *r, &r->xin, np * prm::IN_PAD * 2);
    if (!rc) rc = alloc(c, *r, &r->act, np * 5 * 64 * 2);
    if (!rc) rc = alloc(c, *r, &r->dz, np * (5 * 64 + 16) * 2);
    if (!rc) rc = alloc(c, *r, &r->partial, (size_t)r->max_chunks * prm::N_PARAMS * 4);
    if (!rc) rc = alloc(c, *r, &r->loss_part, np / 16 * 4);
    if (!rc) rc = alloc(c, *r, &r->ctr, 8 * 4);
    if (!rc) rc = alloc(c, *r, &r->lossv, 2 * 4);
    if (rc) {
        free_cache_rec(r);
        return rc;
    }
    // Uniform(+-sqrt(6 / (fan_in + fan_out))) on the used entries; the 20
    // always-zero input channels of layer 0 and output rows 3.. of layer 5
    // stay zero (they receive no gradient).
    std::vector<float> w(prm::N_PARAMS, 0.f);
    SplitMix rng{seed};
    for (int L = 0; L < prm::N_LAYERS; L++) {
        const prm::Layer Y = prm::layer(L);
        const int rows = L == 5 ? 3 : Y.out, used_in = L == 0 ? 172 : Y.in;
        const double a = std::sqrt(6.0 / (double)(used_in + rows));
        for (int o = 0; o < rows; o++)
            for (int k = 0; k < Y.in; k++)
                if (L != 0 || prm::input_used(k)) w[Y.pofs + o * Y.in + k] = (float)((2.0 * rng.u01() - 1.0) * a);
    }
    hipError_t e = hipMemcpy(r->master, w.data(), prm::N_PARAMS * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(r->partial, 0, (size_t)r->max_chunks * prm::N_PARAMS * 4);
    if (e == hipSuccess) e = hipMemset(r->ffrag, 0, prm::N_FFRAG * 64 * 16);
    if (e == hipSuccess) e = hipMemset(r->bfrag, 0, prm::N_BFRAG * 64 * 16);
    if (e != hipSuccess) {
        free_cache_rec(r);
        return fail(c, PRAY_ERR_HIP, "cache set-up failed: %s", hipGetErrorString(e));
    }
    rc = reset_optimiser(c, *r);
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(c, PRAY_ERR_HIP, "cache set-up failed");
    if (rc) {
        free_cache_rec(r);
        return rc;
    }
    const int id = c->next_id++;
    c->caches[id] = r;
    *cache_id = id;
    return PRAY_OK;
}

int pray_cache_destroy(pray_ctx *c, int id) {
    if (!c) return PRAY_ERR_INVALID;
    CacheRec *r = find_cache(c, id);
    if (!r) return fail(c, PRAY_ERR_NOSCENE, "unknown cache %d", id);
    HIPC(c, hipStreamSynchronize(c->stream));
    free_cache_rec(r);
    c->caches.erase(id);
    return PRAY_OK;
}

int pray_cache_get_params(pray_ctx *c, int id, float *params) {
    if (!c || !params) return fail(c, PRAY_ERR_INVALID, "null argument");
    CacheRec *r = find_cache(c, id);
    if (!r) return fail(c, PRAY_ERR_NOSCENE, "unknown cache %d", id);
    HIPC(c, hipStreamSynchronize(c->stream));
    HIPC(c, hipMemcpy(params, r->master, prm::N_PARAMS * 4, hipMemcpyDeviceToHost));
    return PRAY_OK;
}

int pray_cache_set_params(pray_ctx *c, int id, const float *params) {
    if (!c || !params) return fail(c, PRAY_ERR_INVALID, "null argument");
    CacheRec *r = find_cache(c, id);
    if (!r) return fail(c, PRAY_ERR_NOSCENE, "unknown cache %d", id);
    HIPC(c, hipStreamSynchronize(c->stream));
    HIPC(c, hipMemcpy(r->master, params, prm::N_PARAMS * 4, hipMemcpyHostToDevice));
    int rc = reset_optimiser(c, *r);
    if (rc) return rc;
    HIPC(c, hipStreamSynchronize(c->stream));
    return PRAY_OK;
}

int pray_cache_infer(pray_ctx *c, int id, const float *feats, uint64_t n, float *out) {
    if (!c || (n && (!feats || !out))) return fail(c, PRAY_ERR_INVALID, "null argument");
    CacheRec *r = find_cache(c, id);
    if (!r) return fail(c, PRAY_ERR_NOSCENE, "unknown cache %d", id);
    if (n == 0) return PRAY_OK;
    if (n > (1ull << 28)) return fail(c, PRAY_ERR_INVALID, "too many queries");
    std::vector<float> q(n * 16, 0.f);
    for (uint64_t i = 0; i < n; i++) std::memcpy(&q[16 * i], feats + prm::N_FEAT * i, prm::N_FEAT * 4);
    DBuf dq, dout;
    HIPC(c, hipMalloc(&dq.p, n * 64));
    HIPC(c, hipMalloc(&dout.p, n * 12));
    hipStream_t st = c->stream;
    HIPC(c, hipMemcpyAsync(dq.p, q.data(), n * 64, hipMemcpyHostToDevice, st));
    prm::InferArgs ia;
    ia.queries = (const float *)dq.p;
    ia.n = (uint32_t)n;
    ia.n_dev = nullptr;
    ia.out = (float *)dout.p;
    ia.contrib = nullptr;
    ia.targets = nullptr;
    const unsigned grid = (unsigned)std::min<uint64_t>((n + 63) / 64, (uint64_t)c->num_cus * 8);
    hipLaunchKernelGGL(prm::k_mlp_infer, dim3(grid), dim3(256), 0, st, (const prm::bf16x8 *)r->ffrag, ia);
    HIPC(c, hipGetLastError());
    HIPC(c, hipMemcpyAsync(out, dout.p, n * 12, hipMemcpyDeviceToHost, st));
    HIPC(c, hipStreamSynchronize(st));
    return PRAY_OK;
}

int pray_cache_gradient(pray_ctx *c, int id, const float *feats, const float *targets, uint64_t n, float *grad,
                        float *loss) {
    if (!c || !feats || !targets || !grad) return fail(c, PRAY_ERR_INVALID, "null argument");
    CacheRec *r = find_cache(c, id);
    if (!r) return fail(c, PRAY_ERR_NOSCENE, "unknown cache %d", id);
    HIPC(c, hipStreamSynchronize(c->stream));
    int rc = load_batch(c, *r, feats, targets, n);
    if (!rc) rc = enqueue_train(c, *r, (uint32_t)n, false, 0.f, 0u, nullptr, 0u, 0u, c->stream);
    if (rc) return rc;
    HIPC(c, hipStreamSynchronize(c->stream));
    HIPC(c, hipMemcpy(grad, r->grad, prm::N_PARAMS * 4, hipMemcpyDeviceToHost));
    return batch_loss(c, *r, n, loss);
}

int pray_cache_train_step(pray_ctx *c, int id, const float *feats, const float *targets, uint64_t n, float *loss) {
    if (!c || !feats || !targets) return fail(c, PRAY_ERR_INVALID, "null argument");
    CacheRec *r = find_cache(c, id);
    if (!r) return fail(c, PRAY_ERR_NOSCENE, "unknown cache %d", id);
    HIPC(c, hipStreamSynchronize(c->stream));
    int rc = load_batch(c, *r, feats, targets, n);
    if (!rc) rc = enqueue_train(c, *r, (uint32_t)n, true, 0.f, 0xFFFFFFFFu, nullptr, 0u, 0u, c->stream);
    if (rc) return rc;
    HIPC(c, hipStreamSynchronize(c->stream));
    r->steps++;
    return batch_loss(c, *r, n, loss);
}

int pray_cache_stats(pray_ctx *c, int id, uint32_t *steps, float *loss_last, float *loss_running, int *cache_on) {
    if (!c) return PRAY_ERR_INVALID;
    CacheRec *r = find_cache(c, id);
    if (!r) return fail(c, PRAY_ERR_NOSCENE, "unknown cache %d", id);
    HIPC(c, hipStreamSynchronize(c->stream));
    uint32_t ctr[3];
    float lv[2];
    HIPC(c, hipMemcpy(ctr, r->ctr, 12, hipMemcpyDeviceToHost));
    HIPC(c, hipMemcpy(lv, r->lossv, 8, hipMemcpyDeviceToHost));
    if (steps) *steps = ctr[1];
    if (loss_last) *loss_last = lv[0];
    if (loss_running) *loss_running = lv[1];
    if (cache_on) *cache_on = (int)ctr[2];
    return PRAY_OK;
}

int pray_render_pt_cache(pray_ctx *c, int scene_id, int cache_id, const float *cam16, uint32_t w, uint32_t h,
                         uint32_t frames, uint32_t max_depth, uint32_t rr_depth, uint64_t seed, const float *bg,
                         float det_eps, uint32_t train_paths, uint32_t cache_bounce, float max_loss,
                         uint32_t min_steps, float *hdr, uint8_t *rgb, float *frame_stats) {
    return pray_render_pt_cache_ex(c, scene_id, cache_id, cam16, w, h, frames, 1, max_depth, rr_depth, seed, bg,
                                   det_eps, train_paths, cache_bounce, max_loss, min_steps, hdr, rgb, frame_stats);
}

int pray_render_pt_cache_ex(pray_ctx *c, int scene_id, int cache_id, const float *cam16, uint32_t w, uint32_t h,
                            uint32_t frames, uint32_t spf, uint32_t max_depth, uint32_t rr_depth, uint64_t seed,
                            const float *bg, float det_eps, uint32_t train_paths, uint32_t cache_bounce,
                            float max_loss, uint32_t min_steps, float *hdr, uint8_t *rgb, float *frame_stats) {
    if (!c) return PRAY_ERR_INVALID;
    SceneRec *sc = find_scene(c, scene_id);
    if (!sc) return fail(c, PRAY_ERR_NOSCENE, "unknown scene %d", scene_id);
    CacheRec *r = find_cache(c, cache_id);
    if (!r) return fail(c, PRAY_ERR_NOSCENE, "unknown cache %d", cache_id);
    if (!cam16 || !bg) return fail(c, PRAY_ERR_INVALID, "null argument");
    const uint64_t npix = (uint64_t)w * h;
    if (npix == 0 || frames == 0) return PRAY_OK;
    if (npix > (1u << 28)) return fail(c, PRAY_ERR_INVALID, "image too large");
    if (spf == 0 || npix * spf > (1ull << 31))
        return fail(c, PRAY_ERR_INVALID, "samples_per_frame x pixels must be in [1, 2^31]");
    if ((uint64_t)frames * spf > (1u << 24)) return fail(c, PRAY_ERR_INVALID, "more than 2^24 samples");
    if (train_paths == 0 || train_paths > r->max_batch)
        return fail(c, PRAY_ERR_INVALID, "train_paths must be in [1, %u]", r->max_batch);
    if (cache_bounce >= max_depth) return fail(c, PRAY_ERR_INVALID, "cache_bounce must be below max_depth");
    if (sc->d.n_tris == 0) return fail(c, PRAY_ERR_INVALID, "the cache path tracer needs a non-empty scene");
    if (!sc->d.normal || !sc->d.color || !sc->d.emit)
        return fail(c, PRAY_ERR_INVALID, "scene was uploaded without shading data");
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, hipStreamSynchronize(c->stream));
    const uint32_t n_train = (uint32_t)std::min<uint64_t>(train_paths, npix);
    const uint32_t stride = (uint32_t)std::max<uint64_t>(1, npix / n_train);
    prr::Cam cam;
    std::memcpy(&cam, cam16, sizeof cam);
    // Buffers from the context's render workspace (slots 16..36), reused by
    // later calls.
    struct Buf {
        void *p = nullptr;
    } prim, rq[2], sq[2], cnt, hit, tt, contrib, sum, outh, outb, queries, hist, trays, tstate, thit, ttt, tq, phit,
        pt;
    const uint64_t npath = npix * spf;  // paths of a frame: spf samples of every pixel
    WS_GET(c, 16, prim.p, npix * 24);
    for (int q = 0; q < 2; q++) {
        WS_GET(c, 17 + q, rq[q].p, npath * 24);
        WS_GET(c, 19 + q, sq[q].p, npath * 16);
    }
    // counters: [0], [1] path queues, [2] first sample of the frame,
    // [3] render queries, [4] training queries
    WS_GET(c, 21, cnt.p, 8 * sizeof(uint32_t));
    WS_GET(c, 22, hit.p, npath * 4);
    WS_GET(c, 23, tt.p, npath * 4);
    WS_GET(c, 24, contrib.p, npath * 12);
    WS_GET(c, 25, sum.p, npix * 12);
    WS_GET(c, 26, outh.p, npix * 12);
    WS_GET(c, 27, queries.p, npath * 64);
    WS_GET(c, 28, hist.p, (size_t)frames * 8);
    WS_GET(c, 29, trays.p, (size_t)n_train * 24);
    WS_GET(c, 30, tstate.p, (size_t)n_train * 16);
    WS_GET(c, 31, thit.p, (size_t)n_train * 4);
    WS_GET(c, 32, ttt.p, (size_t)n_train * 4);
    WS_GET(c, 33, tq.p, (size_t)n_train * 64);
    if (rgb) WS_GET(c, 34, outb.p, npix * 3);
    WS_GET(c, 35, phit.p, npix * 4);  // the camera rays' hits, shared by all samples
    WS_GET(c, 36, pt.p, npix * 4);
    uint32_t *count = (uint32_t *)cnt.p;
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(prr::k_camera, dim3(nblocks(npix, 256)), dim3(256), 0, st, cam, w, h, (float *)prim.p);
    hipLaunchKernelGGL(prr::k_fill_f32, dim3(nblocks(3 * npix, 256)), dim3(256), 0, st, (float *)sum.p,
                       (uint32_t)(3 * npix), 0.f);
    hipLaunchKernelGGL(prr::k_fill_f32, dim3(nblocks(2 * (uint64_t)frames, 256)), dim3(256), 0, st, (float *)hist.p,
                       2u * frames, -1.f);
    hipLaunchKernelGGL(prt::k_zero_u32, dim3(1), dim3(64), 0, st, count, 8u);
    HIPC(c, hipGetLastError());
    prr::PtArgs P;
    P.n_pix = (uint32_t)npix;
    P.max_depth = max_depth;
    P.rr_depth = rr_depth;
    P.k0 = (uint32_t)seed, P.k1 = (uint32_t)(seed >> 32);
    P.bg[0] = bg[0], P.bg[1] = bg[1], P.bg[2] = bg[2];
    P.sample = count + 2;
    P.samples = spf;
    prr::CacheArgs C{};
    C.cache_bounce = cache_bounce;
    C.cache_on = r->ctr + 2;
    C.queries = (float *)queries.p;
    C.n_queries = count + 3;
    C.n_train = n_train;
    C.stride = stride;
    C.feats = r->feats;
    C.train_rays = (float *)trays.p;
    C.train_state = (float4 *)tstate.p;
    C.n_train_rays = r->ctr;  // also the number of valid training samples
    for (int k = 0; k < 3; k++) {
        const float lo = sc->d.box[k], ext = sc->d.box[k + 3] - sc->d.box[k];
        C.lo[k] = lo;
        C.inv[k] = ext > 0.f ? 1.f / ext : 0.f;
    }
    // Every sample of a pixel starts with the pixel's camera ray: its hit is
    // traced once for all frames.
    int rc = enqueue_trace(c, *sc, (const float *)prim.p, npix, nullptr, PRAY_TRACE_CLOSEST, det_eps, nullptr,
                           (int32_t *)phit.p, (float *)pt.p, nullptr, st);
    if (rc) return rc;
    rc = prepare_trace(c, *sc, std::max<uint64_t>(npath, n_train));
    if (rc) return rc;
    if (!c->side) HIPC(c, hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    if (!c->fork_ev) HIPC(c, hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming));
    if (!c->join_ev) HIPC(c, hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming));
    hipStream_t side = c->side;
    const uint32_t step_base = r->steps;
    const unsigned infer_grid = (unsigned)std::min<uint64_t>((npath + 63) / 64, c->num_cus * 8ull);

    // One frame (one sample per pixel): paths end in the cache at bounce
    // cache_bounce while it is on; training pixels leave training rays; the
    // cache predicts the pixels' and the training samples' remaining
    // radiance; one training step; accumulation.
    auto frame = [&]() -> int {
        Range rf("cache frame");
        hipLaunchKernelGGL(prr::k_fill_f32, dim3(nblocks(4 * (uint64_t)n_train, 256)), dim3(256), 0, st,
                           (float *)r->targets, 4u * n_train, 0.f);
        hipLaunchKernelGGL(prt::k_zero_u32, dim3(1), dim3(64), 0, st, r->ctr, 1u);
        hipLaunchKernelGGL(prr::k_pt_generate, dim3(nblocks(npath, 256)), dim3(256), 0, st, (uint32_t)npath,
                           (float4 *)sq[0].p, count + 0);
        hipLaunchKernelGGL(prt::k_zero_u32, dim3(1), dim3(64), 0, st, count + 3, 2u);
        for (uint32_t b = 0; b <= max_depth; b++) {
            const int cur = b & 1, nxt = cur ^ 1;
            hipLaunchKernelGGL(prt::k_zero_u32, dim3(1), dim3(64), 0, st, count + nxt, 1u);
            const bool first_bounce = b == 0;  // hits of the camera rays, traced once
            if (!first_bounce) {
                Range r("cache extend");
                const int e = enqueue_trace(c, *sc, (const float *)rq[cur].p, npath, count + cur, PRAY_TRACE_CLOSEST,
                                            det_eps, nullptr, (int32_t *)hit.p, (float *)tt.p, nullptr, st);
                if (e) return e;
            }
            Range r("cache shade");
            hipLaunchKernelGGL(prr::k_pt_shade<prr::PT_CACHE>, dim3(nblocks(npath, 256)), dim3(256), 0, st, sc->d, P,
                               C, b, first_bounce ? 1 : 0, (const float *)(first_bounce ? prim.p : rq[cur].p),
                               (const float4 *)sq[cur].p, (const uint32_t *)(count + cur),
                               (const int32_t *)(first_bounce ? phit.p : hit.p),
                               (const float *)(first_bounce ? pt.p : tt.p), (float *)rq[nxt].p, (float4 *)sq[nxt].p,
                               count + nxt, (float *)contrib.p, prr::BatchArgs{});
        }
        // The training rays (few, latency bound) and their targets run on the
        // side stream while the pixels' inference runs on the main one.
        int e;
        {
            Range ri("cache training targets and inference");
            HIPC(c, hipEventRecord(c->fork_ev, st));
            HIPC(c, hipStreamWaitEvent(side, c->fork_ev, 0));
            e = enqueue_trace(c, *sc, (const float *)trays.p, n_train, r->ctr, PRAY_TRACE_CLOSEST, det_eps, nullptr,
                              (int32_t *)thit.p, (float *)ttt.p, nullptr, side);
            if (e) return e;
            hipLaunchKernelGGL(prr::k_pt_train_targets, dim3(nblocks(n_train, 256)), dim3(256), 0, side, sc->d, P, C,
                               (const float *)trays.p, (const float4 *)tstate.p, (const uint32_t *)r->ctr,
                               (const int32_t *)thit.p, (const float *)ttt.p, r->targets, (float *)tq.p, count + 4);
            prm::InferArgs it{};
            it.queries = (const float *)tq.p;
            it.n = n_train;
            it.n_dev = count + 4;
            it.targets = r->targets;
            hipLaunchKernelGGL(prm::k_mlp_infer, dim3((unsigned)std::min<uint64_t>((n_train + 63) / 64, infer_grid)),
                               dim3(256), 0, side, (const prm::bf16x8 *)r->ffrag, it);
            HIPC(c, hipEventRecord(c->join_ev, side));
            prm::InferArgs ia{};
            ia.queries = (const float *)queries.p;
            ia.n = (uint32_t)npath;
            ia.n_dev = count + 3;
            ia.contrib = (float *)contrib.p;
            hipLaunchKernelGGL(prm::k_mlp_infer, dim3(infer_grid), dim3(256), 0, st, (const prm::bf16x8 *)r->ffrag, ia);
            HIPC(c, hipStreamWaitEvent(st, c->join_ev, 0));
        }
        {
            Range rt("cache train");
            e = enqueue_train(c, *r, n_train, true, max_loss, min_steps, (float *)hist.p, step_base, frames, st);
        }
        if (e) return e;
        hipLaunchKernelGGL(prr::k_pt_accumulate, dim3(nblocks(3 * npix, 256)), dim3(256), 0, st, (uint32_t)npix, spf,
                           (const float *)contrib.p, (float *)sum.p, count + 2);
        HIPC(c, hipGetLastError());
        return PRAY_OK;
    };

    if (c->use_graphs) {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        HIPC(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        rc = frame();
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
        // completes before the first replay: the replays, whose frames fork
        // onto a second stream, are not left to order themselves after it.
        e = hipStreamSynchronize(st);
        for (uint32_t f = 0; f < frames && e == hipSuccess; f++) e = hipGraphLaunch(exec, st);
        const hipError_t e2 = hipStreamSynchronize(st);
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess || e2 != hipSuccess)
            return fail(c, PRAY_ERR_HIP, "graph launch failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    } else {
        // the same frames enqueued kernel by kernel
        for (uint32_t f = 0; f < frames; f++) {
            rc = frame();
            if (rc) return rc;
        }
        HIPC(c, hipStreamSynchronize(st));
    }
    r->steps += frames;
    hipLaunchKernelGGL(prr::k_pt_finalize, dim3(nblocks(3 * npix, 256)), dim3(256), 0, st, (uint32_t)npix,
                       (const float *)sum.p, (float)((uint64_t)frames * spf), (float *)outh.p, (uint8_t *)outb.p);
    HIPC(c, hipGetLastError());
    if (hdr) HIPC(c, hipMemcpyAsync(hdr, outh.p, npix * 12, hipMemcpyDeviceToHost, st));
    if (rgb) HIPC(c, hipMemcpyAsync(rgb, outb.p, npix * 3, hipMemcpyDeviceToHost, st));
    if (frame_stats) HIPC(c, hipMemcpyAsync(frame_stats, hist.p, (size_t)frames * 8, hipMemcpyDeviceToHost, st));
    HIPC(c, hipStreamSynchronize(st));
    return PRAY_OK;
}

}  // extern "C"
