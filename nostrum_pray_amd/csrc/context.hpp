// Host-side state of a pray_ctx shared by the library's translation units
// (tracer and radiance cache).  Internal: none of this is exported.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pray_hip.h"
#include "traverse.hpp"

// A roctx range for the lifetime of a scope: tools that record roctx
// markers (rocprofv3 --marker-trace) show when the host enqueued each stage
// of a render (the kernels run later on the stream, or inside a graph
// replay).  Without such a tool a range costs a call into an empty stub.
struct Range {
    explicit Range(const char *name) { roctxRangePushA(name); }
    ~Range() { roctxRangePop(); }
    Range(const Range &) = delete;
    Range &operator=(const Range &) = delete;
};

namespace prc {

struct Tuning {
    int layout = PRAY_LAYOUT_PACKED;
    int stack_lds = 6;
    int leaf_ilp = 2;
    int fetch_min = 16;
    int occupancy = 100;  // percent of the resident persistent waves launched
    int wg_waves = 1;     // waves per workgroup (1, 2 or 4)
    int leaf_batch = 0;   // 0: node and leaf steps in every iteration; T: leaf steps once T lanes wait at leaves
    int node_levels = 1;  // node records: 1 children pairs, 2 two-level records (SAH trees), 3 boxed children
                          // pairs (median trees)
};

struct SceneRec {
    prt::DevScene d;
    std::vector<void *> allocs;
    Tuning tune;
};

struct CacheRec;  // radiance cache (cache.hip)
void free_cache_rec(CacheRec *r);

}  // namespace prc

struct pray_ctx {
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::map<int, std::unique_ptr<prc::SceneRec>> scenes;
    std::map<int, prc::CacheRec *> caches;
    int next_id = 1;
    hipEvent_t events[16] = {};   // pray_event_record slots
    hipEvent_t ievents[4] = {};   // internal timing (autotuner, batch)
    // pray_kernel_timing: event pairs around the traversal kernels of the
    // last KT_RING timed traces, a ring indexed by kt_total % KT_RING
    static constexpr int KT_RING = 64;
    hipEvent_t kevents[2 * KT_RING] = {};
    bool kernel_timing = false;
    uint64_t kt_total = 0;        // traversal kernels timed since timing was enabled
    hipStream_t side = nullptr;   // second stream for work overlapped inside graphs
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    bool use_graphs = true;       // path tracers replay captured hipGraphs (pray_use_graphs)
    uint64_t path_budget = 1ull << 24;  // paths in flight in pray_render_pt* (pray_set_path_budget)
    uint64_t last_rays = 0;       // rays traced by the last render call
    std::vector<uint64_t> last_queues;  // per bounce: paths traced (pray_render_pt*)
    uint32_t *work = nullptr;   // ray counters of the persistent kernels
    uint4 *spill = nullptr;     // per-lane stack spill area
    void *scratch = nullptr;    // device address of this device's static scratch (pray_trace_device_noalloc)
    size_t spill_bytes = 0;
    std::map<std::pair<const void *, int>, int> occupancy;  // (kernel, waves per block) -> resident blocks per CU
    std::map<int, prc::Tuning> tuned;        // scene-size bucket -> autotuned tuning
    std::shared_ptr<void> build_scratch;     // device buffers of the tree builds, kept across builds
    // Device buffers of the renderers by slot, kept across calls and grown on
    // demand (render calls return after their work is done, so no buffer is
    // in use when a later call reuses or replaces it).
    std::vector<void *> ws_ptr;
    std::vector<size_t> ws_bytes;
};

namespace prc {

int fail(pray_ctx *c, int code, const char *fmt, ...);

#define HIPC(ctx, expr)                                                                                       \
    do {                                                                                                      \
        hipError_t e_ = (expr);                                                                               \
        if (e_ != hipSuccess)                                                                                 \
            return ::prc::fail((ctx), e_ == hipErrorOutOfMemory ? PRAY_ERR_NOMEM : PRAY_ERR_HIP,            \
                               "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__);          \
    } while (0)

inline hipStream_t pick(pray_ctx *c, void *stream) { return stream ? (hipStream_t)stream : c->stream; }

inline unsigned nblocks(uint64_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

struct DBuf {
    void *p = nullptr;
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
};

// Deepest tree the traversal's spill area is sized for.
constexpr uint32_t MAX_TREE_DEPTH = 64;
// Most stack entries of boxed node steps: the binary layout's depth (the
// tree's plus the splits of its leaves), or three per level of its 4-wide
// form.
constexpr uint32_t MAX_BOXED_DEPTH = 192;

SceneRec *find_scene(pray_ctx *c, int id);
// Device allocation owned by a scene record (freed by free_scene_rec).
int device_alloc(pray_ctx *c, SceneRec &sc, void **p, size_t bytes);
void free_scene_rec(SceneRec &sc);
// Builds the scene's two-level node records (prt::k_build_quads) from its
// uploaded children pairs, synchronising on st, and keeps them in d.quads when
// the tree is in SAH form (every pair's two clips equal) and its indices fit;
// otherwise d.quads stays null and only one-level node steps trace the scene.
int build_quads(pray_ctx *c, SceneRec &sc, hipStream_t st);
int ensure_spill(pray_ctx *c, size_t bytes);
// Render workspace slot `slot` with at least `bytes` bytes (kept by the
// context until pray_destroy).  Slots: render_pt 0..15, the cache renderer
// 16..39, the batch renderer 40..63, render_rt 64..71.
int ws_get(pray_ctx *c, unsigned slot, size_t bytes, void **out);
#define WS_GET(ctx, slot, ptr, bytes)                                                                        \
    do {                                                                                                      \
        void *ws_p_ = nullptr;                                                                                \
        const int ws_rc_ = ::prc::ws_get((ctx), (slot), (bytes), &ws_p_);                                     \
        if (ws_rc_) return ws_rc_;                                                                            \
        (ptr) = ws_p_;                                                                                        \
    } while (0)
// {layout, stack_lds, leaf_ilp, fetch_min, occupancy, wg_waves, leaf_batch, node_levels}
void tuning_values(const Tuning &t, int *v);
// Enqueues one trace over n rays (or *n_dev rays when n_dev is set, with n as
// the launch bound).
int enqueue_trace(pray_ctx *c, const SceneRec &sc, const float *rays, uint64_t n, const uint32_t *n_dev, int mode,
                  float eps, const float *lim2, int32_t *idx, float *t, uint32_t *counts, hipStream_t st);
// Makes sure the closest-hit kernel of a scene has its stack spill area for
// launches of up to n rays (needed before a graph capture).
int prepare_trace(pray_ctx *c, const SceneRec &sc, uint64_t n);
// Closest hits of rays from several scenes (packed layout): ray i belongs to
// table[ray_scene[i]]; max_depth is the deepest tree of the table.
int enqueue_trace_multi(pray_ctx *c, const prt::DevScene *table, const uint32_t *ray_scene, uint32_t max_depth,
                        const Tuning &tu, const float *rays, uint64_t n, const uint32_t *n_dev, float eps,
                        int32_t *idx, float *t, hipStream_t st);
int prepare_trace_multi(pray_ctx *c, const Tuning &tu, uint32_t max_depth, uint64_t n);
// Scene-size bucket of the autotuner (factors of 4 in triangle count).
int size_bucket(uint32_t n_tris);

}  // namespace prc
