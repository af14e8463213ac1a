// libpray_hip.so: gfx950 tracer behind the C ABI of include/pray_hip.h.
//
// Device layout of a scene (uploaded once, resident in HBM):
//   nodes    uint2[n_nodes]          8-byte nodes (format in host/gpu_layout.hpp)
//   refs     uint32[n_refs + 2]      leaf triangle-index runs, bit 31 = last
//   a,ab,ac  float[3*n_tris] each    triangle vertex a and edges (per-field arrays)
//   recs     float4[3*(n_refs + 2)]  packed per-leaf records {a, idx|last}, {ab, 0}, {ac, 0}
//   normal, color float[3*n_tris], emit uint8[n_tris]   shading data (optional)
// The two trailing pad entries keep the leaf look-ahead loads in bounds.
// Traversal kernels: traverse.hpp.  Renderers: render.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pray_hip.h"
#include "context.hpp"
#include "device_math.hpp"
#include "render.hpp"
#include "traverse.hpp"

using prt::DevScene;
using prt::TraceArgs;
using prt::WAVE;

namespace {

// Packed per-leaf records from the per-field arrays.
__global__ void k_pack_records(const uint32_t *refs, uint32_t n_refs, const float *a, const float *ab,
                               const float *ac, float4 *recs) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_refs) return;
    const uint32_t ref = refs[k];
    const size_t t = ref & ~prt::REF_LAST;
    recs[3 * (size_t)k] = make_float4(a[3 * t], a[3 * t + 1], a[3 * t + 2], __uint_as_float(ref));
    recs[3 * (size_t)k + 1] = make_float4(ab[3 * t], ab[3 * t + 1], ab[3 * t + 2], 0.f);
    recs[3 * (size_t)k + 2] = make_float4(ac[3 * t], ac[3 * t + 1], ac[3 * t + 2], 0.f);
}

__global__ void k_fill_miss(uint32_t n, int32_t *out_idx, float *out_t, uint32_t *counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out_idx[i] = -1;
    if (out_t) out_t[i] = FLT_MAX;
    if (counts) counts[2 * (size_t)i] = counts[2 * (size_t)i + 1] = 0;
}

// Incoherent ray stream (see pray_gen_rays_device).
__global__ void k_gen_rays(uint32_t k0, uint32_t k1, float lx, float ly, float lz, float sx, float sy, float sz,
                           uint32_t n, float *rays) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4 r = prd::philox4x32_10(make_uint4(i, 0u, 0u, 0u), k0, k1);
    const uint4 q = prd::philox4x32_10(make_uint4(i, 1u, 0u, 0u), k0, k1);
    const float ox = lx + sx * prd::u01(r.x);
    const float oy = ly + sy * prd::u01(r.y);
    const float oz = lz + sz * prd::u01(r.z);
    const float z = 1.0f - 2.0f * prd::u01(r.w);
    const float rxy = sqrtf(fmaxf(0.f, 1.0f - z * z));
    float sp, cp;
    prd::sincos2pi(prd::u01(q.x), sp, cp);
    float dx = rxy * cp, dy = rxy * sp, dz = z;
    prd::normalize3(dx, dy, dz);
    float *o = rays + 6 * (size_t)i;
    o[0] = ox, o[1] = oy, o[2] = oz, o[3] = dx, o[4] = dy, o[5] = dz;
}

using prr::Cam;

// Grid-stride streaming copy (one 16-byte load and store per element).
__global__ void k_copy16(uint4 *__restrict__ dst, const uint4 *__restrict__ src, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

// Scratch of pray_trace_device_noalloc: the NCTR chunk counters of the
// persistent trace kernels (first NOALLOC_COUNTER_BYTES), then their per-lane
// stack spill area.  Every GPU the library runs on holds its own copy from
// the code object's load until the process ends; all calls on one GPU share
// it.  Host code never passes this name to a kernel or a copy: it is a
// host-side stand-in, and the device address comes from hipGetSymbolAddress
// (scratch_setup).
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
constexpr size_t NOALLOC_COUNTER_BYTES = (size_t)prt::NCTR * prt::CTR_STRIDE * sizeof(uint32_t);
static_assert(NOALLOC_COUNTER_BYTES % 256 == 0, "spill area stays 256-byte aligned");
static __device__ __align__(256) uint4 scratch[NOALLOC_SCRATCH_BYTES / sizeof(uint4)];

namespace {

std::mutex scratch_mu;
std::map<int, void *> scratch_addr;  // device -> address of that device's `scratch`

// Looks the device address of `scratch` up once per device, with `device`
// current (the caller has made it current).
hipError_t scratch_setup(int device, void **out) {
    std::lock_guard<std::mutex> lock(scratch_mu);
    auto it = scratch_addr.find(device);
    if (it != scratch_addr.end()) {
        *out = it->second;
        return hipSuccess;
    }
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return e != hipSuccess ? e : hipErrorInvalidSymbol;
    scratch_addr[device] = p;
    *out = p;
    return hipSuccess;
}

}  // namespace

namespace prc {

int fail(pray_ctx *c, int code, const char *fmt, ...) {
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return code;
}

typedef void (*TraceKernel)(DevScene, TraceArgs);

// LDS ring sizes: 12 bytes per entry and lane, so a 64-lane block takes
// 768*K bytes; with 160 KB of LDS per CU, K = 6 is the largest ring that
// leaves room for 32 resident waves (the VGPR limit of the ILP-2 kernels)
// and K = 5 for 36 (the ILP-1 kernels).
template <bool ANY, int LAYOUT, bool BATCH, int BOX = 0>
TraceKernel pick_kernel_kl(int K, int ilp) {
    using namespace prt;
    switch (K) {
    case 4: return ilp == 2 ? k_trace<ANY, LAYOUT, 4, 2, false, BATCH, BOX> : k_trace<ANY, LAYOUT, 4, 1, false, BATCH, BOX>;
    case 5: return ilp == 2 ? k_trace<ANY, LAYOUT, 5, 2, false, BATCH, BOX> : k_trace<ANY, LAYOUT, 5, 1, false, BATCH, BOX>;
    case 6: return ilp == 2 ? k_trace<ANY, LAYOUT, 6, 2, false, BATCH, BOX> : k_trace<ANY, LAYOUT, 6, 1, false, BATCH, BOX>;
    case 16:
        return ilp == 2 ? k_trace<ANY, LAYOUT, 16, 2, false, BATCH, BOX> : k_trace<ANY, LAYOUT, 16, 1, false, BATCH, BOX>;
    default: return ilp == 2 ? k_trace<ANY, LAYOUT, 8, 2, false, BATCH, BOX> : k_trace<ANY, LAYOUT, 8, 1, false, BATCH, BOX>;
    }
}

TraceKernel pick_kernel_multi(int K, int ilp) {
    switch (K) {
    case 4: return ilp == 2 ? prt::k_trace<false, 1, 4, 2, true> : prt::k_trace<false, 1, 4, 1, true>;
    case 5: return ilp == 2 ? prt::k_trace<false, 1, 5, 2, true> : prt::k_trace<false, 1, 5, 1, true>;
    case 6: return ilp == 2 ? prt::k_trace<false, 1, 6, 2, true> : prt::k_trace<false, 1, 6, 1, true>;
    case 16: return ilp == 2 ? prt::k_trace<false, 1, 16, 2, true> : prt::k_trace<false, 1, 16, 1, true>;
    default: return ilp == 2 ? prt::k_trace<false, 1, 8, 2, true> : prt::k_trace<false, 1, 8, 1, true>;
    }
}

template <bool ANY, bool BATCH>
TraceKernel pick_kernel_q(int K, int ilp) {
    switch (K) {
    case 4: return ilp == 2 ? prt::k_trace_q<ANY, 4, 2, BATCH> : prt::k_trace_q<ANY, 4, 1, BATCH>;
    case 5: return ilp == 2 ? prt::k_trace_q<ANY, 5, 2, BATCH> : prt::k_trace_q<ANY, 5, 1, BATCH>;
    case 6: return ilp == 2 ? prt::k_trace_q<ANY, 6, 2, BATCH> : prt::k_trace_q<ANY, 6, 1, BATCH>;
    case 16: return ilp == 2 ? prt::k_trace_q<ANY, 16, 2, BATCH> : prt::k_trace_q<ANY, 16, 1, BATCH>;
    default: return ilp == 2 ? prt::k_trace_q<ANY, 8, 2, BATCH> : prt::k_trace_q<ANY, 8, 1, BATCH>;
    }
}

int node_form(const SceneRec &sc, const Tuning &tu) {
    if (tu.layout != PRAY_LAYOUT_PACKED) return 1;
    if (tu.node_levels == 2 && sc.d.quads) return 2;
    if (tu.node_levels == 3 && sc.d.boxes) return 3;
    if (tu.node_levels == 4 && sc.d.boxes4) return 4;
    return 1;
}

// Stack entries a lane may need (the tree's depth; the boxed layout's is
// deeper where it splits leaves).
uint32_t trace_depth(const SceneRec &sc, const Tuning &tu) {
    const int f = node_form(sc, tu);
    return std::max<uint32_t>(1, f == 3 ? sc.d.box_depth : (f == 4 ? sc.d.box4_depth : sc.d.max_depth));
}

// The kernel of a tuning: form 2 (two-level records) or 3 (boxed children
// pairs) as node_form resolves them; batch for leaf batching (packed layout
// only).
TraceKernel pick_kernel(bool any, int layout, int K, int ilp, int form, bool batch) {
    if (layout != PRAY_LAYOUT_PACKED)
        return any ? pick_kernel_kl<true, PRAY_LAYOUT_INDEXED, false>(K, ilp)
                   : pick_kernel_kl<false, PRAY_LAYOUT_INDEXED, false>(K, ilp);
    if (form == 2) {
        if (batch) return any ? pick_kernel_q<true, true>(K, ilp) : pick_kernel_q<false, true>(K, ilp);
        return any ? pick_kernel_q<true, false>(K, ilp) : pick_kernel_q<false, false>(K, ilp);
    }
    if (form == 3) {
        if (batch)
            return any ? pick_kernel_kl<true, PRAY_LAYOUT_PACKED, true, 1>(K, ilp)
                       : pick_kernel_kl<false, PRAY_LAYOUT_PACKED, true, 1>(K, ilp);
        return any ? pick_kernel_kl<true, PRAY_LAYOUT_PACKED, false, 1>(K, ilp)
                   : pick_kernel_kl<false, PRAY_LAYOUT_PACKED, false, 1>(K, ilp);
    }
    if (form == 4) {
        if (batch)
            return any ? pick_kernel_kl<true, PRAY_LAYOUT_PACKED, true, 2>(K, ilp)
                       : pick_kernel_kl<false, PRAY_LAYOUT_PACKED, true, 2>(K, ilp);
        return any ? pick_kernel_kl<true, PRAY_LAYOUT_PACKED, false, 2>(K, ilp)
                   : pick_kernel_kl<false, PRAY_LAYOUT_PACKED, false, 2>(K, ilp);
    }
    if (batch)
        return any ? pick_kernel_kl<true, PRAY_LAYOUT_PACKED, true>(K, ilp)
                   : pick_kernel_kl<false, PRAY_LAYOUT_PACKED, true>(K, ilp);
    return any ? pick_kernel_kl<true, PRAY_LAYOUT_PACKED, false>(K, ilp)
               : pick_kernel_kl<false, PRAY_LAYOUT_PACKED, false>(K, ilp);
}

// Resident blocks per CU of a trace kernel with `wpw` waves per block and
// `lds` bytes of dynamic LDS per block (cached).
int resident_blocks(pray_ctx *c, TraceKernel k, int wpw, size_t lds) {
    const auto key = std::make_pair((const void *)k, wpw);
    auto it = c->occupancy.find(key);
    if (it != c->occupancy.end()) return it->second;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, wpw * WAVE, lds) != hipSuccess || nb <= 0)
        nb = std::max(1, 8 / wpw);
    c->occupancy[key] = nb;
    return nb;
}

// Persistent grid: one wave per 64 rays, at most the tuned share of the
// waves that fit on the device at once, in blocks of tu.wg_waves waves.
struct Launch {
    unsigned blocks = 1, threads = WAVE;
    size_t lds = 0;         // dynamic LDS per block: the waves' stack rings
    uint64_t lanes = WAVE;  // lanes of the grid (stack spill slots)
};

Launch launch_shape(pray_ctx *c, TraceKernel k, const Tuning &tu, uint64_t n) {
    const int wpw = tu.wg_waves;
    Launch L;
    L.threads = (unsigned)(wpw * WAVE);
    L.lds = (size_t)wpw * prt::ring_lds_bytes(tu.stack_lds);
    const uint64_t want = (n + WAVE - 1) / WAVE;
    const uint64_t per_cu = std::max<uint64_t>(
        1, (uint64_t)resident_blocks(c, k, wpw, L.lds) * (uint64_t)wpw * (uint64_t)tu.occupancy / 100);
    const uint64_t waves = std::max<uint64_t>(1, std::min(want, (uint64_t)c->num_cus * per_cu));
    L.blocks = (unsigned)((waves + wpw - 1) / wpw);
    L.lanes = (uint64_t)L.blocks * wpw * WAVE;
    return L;
}

int ensure_spill(pray_ctx *c, size_t bytes) {
    if (bytes <= c->spill_bytes) return PRAY_OK;
    if (c->spill) {
        // traces may run on callers' streams: wait for all of them
        HIPC(c, hipDeviceSynchronize());
        HIPC(c, hipFree(c->spill));
        c->spill = nullptr;
        c->spill_bytes = 0;
    }
    HIPC(c, hipMalloc((void **)&c->spill, bytes));
    c->spill_bytes = bytes;
    return PRAY_OK;
}

// Scratch of a trace whose caller provides it instead of the context: the
// NCTR chunk counters (NCTR * CTR_STRIDE words) and a stack spill area of
// spill_bytes; the persistent grid is cut to the lanes whose spill slots fit.
struct FixedScratch {
    uint32_t *work = nullptr;
    uint4 *spill = nullptr;
    size_t spill_bytes = 0;
};

int ws_get(pray_ctx *c, unsigned slot, size_t bytes, void **out) {
    if (slot >= c->ws_ptr.size()) {
        c->ws_ptr.resize(slot + 1, nullptr);
        c->ws_bytes.resize(slot + 1, 0);
    }
    bytes = std::max<size_t>(bytes, 256);
    if (c->ws_bytes[slot] < bytes) {
        if (c->ws_ptr[slot]) {
            HIPC(c, hipFree(c->ws_ptr[slot]));
            c->ws_ptr[slot] = nullptr;
            c->ws_bytes[slot] = 0;
        }
        HIPC(c, hipMalloc(&c->ws_ptr[slot], bytes));
        c->ws_bytes[slot] = bytes;
    }
    *out = c->ws_ptr[slot];
    return PRAY_OK;
}

// Enqueues one trace over n rays (or *n_dev rays when n_dev is set, with n as
// the launch bound), with the context's counters and spill area (grown on
// demand), or with `fixed` scratch (nothing is allocated).
static int enqueue_trace_ex(pray_ctx *c, const SceneRec &sc, const float *rays, uint64_t n, const uint32_t *n_dev,
                            int mode, float eps, const float *lim2, int32_t *idx, float *t, uint32_t *counts,
                            hipStream_t st, const FixedScratch *fixed) {
    if (n == 0) return PRAY_OK;
    if (n > 0x7FFFFFFFull) return fail(c, PRAY_ERR_INVALID, "at most 2^31-1 rays per call");
    if (mode != PRAY_TRACE_CLOSEST && mode != PRAY_TRACE_ANY && mode != PRAY_TRACE_ANY_INCLUSIVE)
        return fail(c, PRAY_ERR_INVALID, "unknown trace mode %d", mode);
    if (mode != PRAY_TRACE_CLOSEST && !lim2) return fail(c, PRAY_ERR_INVALID, "any-hit modes need limit2");
    if (sc.d.n_tris == 0) {
        if (n_dev) return fail(c, PRAY_ERR_INVALID, "device-counted trace on an empty scene");
        hipLaunchKernelGGL(k_fill_miss, dim3(nblocks(n, 256)), dim3(256), 0, st, (uint32_t)n, idx, t, counts);
        HIPC(c, hipGetLastError());
        return PRAY_OK;
    }
    const Tuning &tu = sc.tune;
    TraceKernel k = pick_kernel(mode != PRAY_TRACE_CLOSEST, tu.layout, tu.stack_lds, tu.leaf_ilp, node_form(sc, tu),
                                tu.leaf_batch > 0);
    Launch L = launch_shape(c, k, tu, n);
    const uint32_t depth = trace_depth(sc, tu);
    uint32_t *work = c->work;
    uint4 *spill = c->spill;
    if (fixed) {
        // Fewer persistent waves when the spill slots of all of them do not
        // fit: each takes more chunks, the results are the same.
        const uint64_t lanes_fit = fixed->spill_bytes / ((uint64_t)depth * sizeof(uint4));
        const uint64_t block_lanes = (uint64_t)L.threads;
        if (lanes_fit < block_lanes)
            return fail(c, PRAY_ERR_UNSUPPORTED, "the scratch holds no workgroup's stack spill at tree depth %u", depth);
        if (L.lanes > lanes_fit) {
            L.blocks = (unsigned)(lanes_fit / block_lanes);
            L.lanes = (uint64_t)L.blocks * block_lanes;
        }
        work = fixed->work;
        spill = fixed->spill;
    } else {
        int rc = ensure_spill(c, (size_t)L.lanes * depth * sizeof(uint4));
        if (rc) return rc;
        spill = c->spill;
    }
    TraceArgs a{};
    a.rays = rays;
    a.n = (uint32_t)n;
    a.n_dev = n_dev;
    a.limit2 = lim2;
    a.eps = eps;
    a.inclusive = mode == PRAY_TRACE_ANY_INCLUSIVE;
    a.out_idx = idx;
    a.out_t = t;
    a.counts = counts;
    a.work = work;
    a.spill = spill;
    a.spill_depth = depth;
    a.fetch_min = (uint32_t)tu.fetch_min;
    a.leaf_batch = (uint32_t)tu.leaf_batch;
    hipLaunchKernelGGL(prt::k_reset_counters, dim3(1), dim3(64), 0, st, work);
    bool timed = false;
    if (c->kernel_timing) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        HIPC(c, hipStreamIsCapturing(st, &cs));
        timed = cs == hipStreamCaptureStatusNone;
    }
    const int kslot = (int)(c->kt_total % pray_ctx::KT_RING);
    if (timed) HIPC(c, hipEventRecord(c->kevents[2 * kslot], st));
    hipLaunchKernelGGL(k, dim3(L.blocks), dim3(L.threads), L.lds, st, sc.d, a);
    if (timed) {
        HIPC(c, hipEventRecord(c->kevents[2 * kslot + 1], st));
        c->kt_total++;
    }
    HIPC(c, hipGetLastError());
    return PRAY_OK;
}

int enqueue_trace(pray_ctx *c, const SceneRec &sc, const float *rays, uint64_t n, const uint32_t *n_dev, int mode,
                  float eps, const float *lim2, int32_t *idx, float *t, uint32_t *counts, hipStream_t st) {
    return enqueue_trace_ex(c, sc, rays, n, n_dev, mode, eps, lim2, idx, t, counts, st, nullptr);
}

int enqueue_trace_multi(pray_ctx *c, const prt::DevScene *table, const uint32_t *ray_scene, uint32_t max_depth,
                        const Tuning &tu, const float *rays, uint64_t n, const uint32_t *n_dev, float eps,
                        int32_t *idx, float *t, hipStream_t st) {
    if (n == 0) return PRAY_OK;
    if (n > 0x7FFFFFFFull) return fail(c, PRAY_ERR_INVALID, "at most 2^31-1 rays per call");
    TraceKernel k = pick_kernel_multi(tu.stack_lds, tu.leaf_ilp);
    const Launch L = launch_shape(c, k, tu, n);
    const uint32_t depth = std::max<uint32_t>(1, max_depth);
    int rc = ensure_spill(c, (size_t)L.lanes * depth * sizeof(uint4));
    if (rc) return rc;
    TraceArgs a{};
    a.rays = rays;
    a.n = (uint32_t)n;
    a.n_dev = n_dev;
    a.eps = eps;
    a.out_idx = idx;
    a.out_t = t;
    a.work = c->work;
    a.spill = c->spill;
    a.spill_depth = depth;
    a.fetch_min = (uint32_t)tu.fetch_min;
    a.leaf_batch = (uint32_t)tu.leaf_batch;
    a.table = table;
    a.ray_scene = ray_scene;
    hipLaunchKernelGGL(prt::k_reset_counters, dim3(1), dim3(64), 0, st, c->work);
    hipLaunchKernelGGL(k, dim3(L.blocks), dim3(L.threads), L.lds, st, prt::DevScene{}, a);
    HIPC(c, hipGetLastError());
    return PRAY_OK;
}

int prepare_trace_multi(pray_ctx *c, const Tuning &tu, uint32_t max_depth, uint64_t n) {
    TraceKernel k = pick_kernel_multi(tu.stack_lds, tu.leaf_ilp);
    return ensure_spill(c, (size_t)launch_shape(c, k, tu, n).lanes * std::max<uint32_t>(1, max_depth) * sizeof(uint4));
}

int size_bucket(uint32_t n_tris) {
    int b = 0;
    while (b < 15 && (n_tris >> (2 * b + 2)) >= 1u) ++b;  // 4^b <= n_tris < 4^(b+1)
    return b;
}

SceneRec *find_scene(pray_ctx *c, int id) {
    auto it = c->scenes.find(id);
    return it == c->scenes.end() ? nullptr : it->second.get();
}

int device_alloc(pray_ctx *c, SceneRec &sc, void **p, size_t bytes) {
    if (bytes == 0) {
        *p = nullptr;
        return PRAY_OK;
    }
    HIPC(c, hipMalloc(p, bytes));
    sc.allocs.push_back(*p);
    return PRAY_OK;
}

void free_scene_rec(SceneRec &sc) {
    for (void *p : sc.allocs) (void)hipFree(p);
    sc.allocs.clear();
}

int build_quads(pray_ctx *c, SceneRec &sc, hipStream_t st) {
    prt::DevScene &d = sc.d;
    d.quads = nullptr;
    if (d.n_tris == 0 || d.n_nodes < 2 || d.n_refs >= (prt::Q_MASK30 >> 2)) return PRAY_OK;
    const uint32_t n_pairs = (d.n_nodes + 1) / 2;
    const size_t bytes = (size_t)n_pairs * 32;
    void *p = nullptr;
    // the records, then one word for the build's flag
    int rc = device_alloc(c, sc, &p, bytes + 16);
    if (rc) return rc;
    uint32_t *flag = reinterpret_cast<uint32_t *>(static_cast<char *>(p) + bytes);
    hipLaunchKernelGGL(prt::k_zero_u32, dim3(1), dim3(64), 0, st, flag, 1u);
    hipLaunchKernelGGL(prt::k_build_quads, dim3(nblocks(n_pairs, 256)), dim3(256), 0, st,
                       reinterpret_cast<const uint4 *>(d.nodes), n_pairs, static_cast<uint4 *>(p), flag);
    HIPC(c, hipGetLastError());
    uint32_t f = ~0u;
    HIPC(c, hipMemcpyAsync(&f, flag, 4, hipMemcpyDeviceToHost, st));
    HIPC(c, hipStreamSynchronize(st));
    if (f == 0) {
        d.quads = static_cast<const uint4 *>(p);
    } else {
        // not a SAH-form tree (or indices too large): one-level node steps only
        sc.allocs.pop_back();
        HIPC(c, hipFree(p));
    }
    return PRAY_OK;
}

// ---------------------------------------------------------- boxed layout --
// The layout of boxed node steps (DevScene::boxes, DevScene::brecs), made
// for every uploaded tree (the default node steps of median trees, an
// object partition whose children's boxes overlap; selectable for SAH
// trees).  It has its own pair and leaf-reference index spaces:
// - Every inner node of the uploaded tree becomes a pair record of its two
//   children (16 words: per axis {child 0 lower, child 1 lower, child 0
//   upper, child 1 upper}, then the children's words); pair 0 holds an empty
//   child 0 and the root as child 1.
// - Every leaf of more than BOX_LEAF references is split further: its
//   triangles are sorted by centroid along the longest axis of their
//   centroids' bounds (ties by triangle index) and halved until no leaf holds
//   more, each split a pair record of its own.  The leaves' runs are written
//   in depth-first order to a reference array of their own, with the last
//   entry of each run flagged.
// - A child's box is the union of the boxes of the triangles its subtree
//   references, with the vertices a, a + ab, a + ac summed in double and the
//   bounds rounded outward to float, so that it holds the triangle the
//   intersection test sees, cut to the child's cell: the region the split
//   planes on the path from the root bound (where a children pair's two
//   clips are equal, child 0 lies at or below that plane on its parent's
//   axis and child 1 at or above it).  A KD tree references a triangle from
//   the leaves on both sides of every plane it crosses, and the two sides
//   are cut at the same float, so each point of the triangle lies in the
//   cut box of a leaf that references it; without the cut, a leaf's box
//   would reach as far as its largest triangle beyond the cell.  A pair
//   with two different clips (a median tree's object partition, each
//   triangle referenced once) cuts nothing.  An empty subtree (or cut) gets
//   {+inf, -inf}.
// The traversal's closest hit does not depend on the tree (it is the
// lexicographic minimum of (t, index) over all triangles), so this finer
// tree gives the same hits as the uploaded one.
namespace {

constexpr uint32_t BOX_LEAF = 2;  // 2 to 8 measured: 2 traced the terrain 3.6% and the spheres scene 14% faster than 4

struct Bounds {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Bounds &o) {
        for (int a = 0; a < 3; ++a) lo[a] = std::min(lo[a], o.lo[a]), hi[a] = std::max(hi[a], o.hi[a]);
    }
    // Cuts this box to `cell`; a box left empty becomes {+inf, -inf}.
    void cut(const Bounds &cell) {
        bool empty = false;
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::max(lo[a], cell.lo[a]), hi[a] = std::min(hi[a], cell.hi[a]);
            empty = empty || !(lo[a] <= hi[a]);
        }
        if (empty) *this = Bounds{};
    }
    static Bounds everywhere() {
        Bounds b;
        for (int a = 0; a < 3; ++a) b.lo[a] = -INFINITY, b.hi[a] = INFINITY;
        return b;
    }
};

struct BoxedBuilder {
    const uint32_t *nodes, *refs;
    const float *A, *AB, *AC;
    std::vector<uint32_t> rec;    // 16 words per pair
    std::vector<uint32_t> brefs;  // leaf runs
    uint32_t depth = 0;

    static float down(double x) {
        const float f = (float)x;
        return (double)f > x ? std::nextafter(f, -INFINITY) : f;
    }
    static float up(double x) {
        const float f = (float)x;
        return (double)f < x ? std::nextafter(f, INFINITY) : f;
    }
    Bounds tri_box(uint32_t t) const {
        Bounds b;
        for (int a = 0; a < 3; ++a) {
            const double v0 = A[3 * (size_t)t + a], v1 = v0 + AB[3 * (size_t)t + a], v2 = v0 + AC[3 * (size_t)t + a];
            b.lo[a] = down(std::min(v0, std::min(v1, v2)));
            b.hi[a] = up(std::max(v0, std::max(v1, v2)));
        }
        return b;
    }
    double centroid(uint32_t t, int a) const {
        return (double)A[3 * (size_t)t + a] + ((double)AB[3 * (size_t)t + a] + (double)AC[3 * (size_t)t + a]) / 3.0;
    }
    uint32_t new_pair() {
        rec.resize(rec.size() + 16, 0u);
        return (uint32_t)(rec.size() / 16 - 1);
    }
    void set_child(uint32_t p, int ch, uint32_t w, const Bounds &b) {
        uint32_t *r = &rec[16 * (size_t)p];
        for (int a = 0; a < 3; ++a) {
            std::memcpy(&r[4 * a + ch], &b.lo[a], 4);
            std::memcpy(&r[4 * a + 2 + ch], &b.hi[a], 4);
        }
        r[12 + ch] = w;
    }
    // A leaf of the n triangles at t in `cell`, split while over BOX_LEAF:
    // its word and box.
    uint32_t leaf(uint32_t *t, uint32_t n, uint32_t d, const Bounds &cell, Bounds &box) {
        box = Bounds{};
        if (n <= BOX_LEAF) {
            depth = std::max(depth, d);
            const uint32_t k = (uint32_t)brefs.size();
            for (uint32_t j = 0; j < n; ++j) {
                brefs.push_back(t[j] | (j + 1 == n ? prt::REF_LAST : 0u));
                box.grow(tri_box(t[j]));
            }
            box.cut(cell);
            return (k << 2) | prt::LEAF_TAG;
        }
        double clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t j = 0; j < n; ++j)
            for (int a = 0; a < 3; ++a) {
                const double c = centroid(t[j], a);
                clo[a] = std::min(clo[a], c), chi[a] = std::max(chi[a], c);
            }
        int ax = 0;
        for (int a = 1; a < 3; ++a)
            if (chi[a] - clo[a] > chi[ax] - clo[ax]) ax = a;
        std::vector<std::pair<double, uint32_t>> key(n);
        for (uint32_t j = 0; j < n; ++j) key[j] = {centroid(t[j], ax), t[j]};
        std::sort(key.begin(), key.end());  // by centroid, then index
        for (uint32_t j = 0; j < n; ++j) t[j] = key[j].second;
        const uint32_t p = new_pair();
        Bounds b0, b1;
        const uint32_t w0 = leaf(t, n / 2, d + 1, cell, b0);
        const uint32_t w1 = leaf(t + n / 2, n - n / 2, d + 1, cell, b1);
        set_child(p, 0, w0, b0);
        set_child(p, 1, w1, b1);
        box = b0;
        box.grow(b1);
        return (2 * p) << 2;
    }
    // Node i of the uploaded tree, whose cell is `cell`: its word in this
    // layout and its box.
    uint32_t node(uint32_t i, uint32_t d, const Bounds &cell, Bounds &box) {
        const uint32_t w = nodes[2 * (size_t)i + 1];
        if ((w & 3u) == prt::LEAF_TAG) {
            uint32_t k = w >> 2;
            if (k == prt::LEAF_EMPTY) {
                box = Bounds{};
                return w;
            }
            std::vector<uint32_t> run;
            while (true) {
                run.push_back(refs[k] & ~prt::REF_LAST);
                if (refs[k] & prt::REF_LAST) break;
                ++k;
            }
            return leaf(run.data(), (uint32_t)run.size(), d, cell, box);
        }
        const uint32_t p = new_pair(), axis = w & 3u, lc = w >> 2;
        Bounds c0 = cell, c1 = cell;
        if (nodes[2 * (size_t)lc] == nodes[2 * (size_t)lc + 2]) {
            float plane;
            std::memcpy(&plane, &nodes[2 * (size_t)lc], 4);
            c0.hi[axis] = std::min(c0.hi[axis], plane);
            c1.lo[axis] = std::max(c1.lo[axis], plane);
        }
        Bounds b0, b1;
        const uint32_t w0 = node(lc, d + 1, c0, b0);
        const uint32_t w1 = node(lc + 1, d + 1, c1, b1);
        set_child(p, 0, w0, b0);
        set_child(p, 1, w1, b1);
        box = b0;
        box.grow(b1);
        return (2 * p) << 2;
    }
};

// Builds the layout of a validated tree.
int make_boxed_layout(pray_ctx *c, const uint32_t *nodes, const uint32_t *refs, const float *A, const float *AB,
                      const float *AC, BoxedBuilder &B) {
    B.nodes = nodes, B.refs = refs, B.A = A, B.AB = AB, B.AC = AC;
    const uint32_t p0 = B.new_pair();
    Bounds root;
    const uint32_t rw = B.node(1, 0, Bounds::everywhere(), root);
    B.set_child(p0, 0, (prt::LEAF_EMPTY << 2) | prt::LEAF_TAG, Bounds{});
    B.set_child(p0, 1, rw, root);
    if (B.rec.size() / 16 >= (1u << 29) || B.brefs.size() >= prt::LEAF_EMPTY)
        return fail(c, PRAY_ERR_UNSUPPORTED, "tree too large for the boxed layout");
    if (B.depth > MAX_BOXED_DEPTH) return fail(c, PRAY_ERR_UNSUPPORTED, "boxed layout deeper than %u", MAX_BOXED_DEPTH);
    return PRAY_OK;
}

// The 4-wide form of a boxed layout (node_levels 4): each record collapses a
// binary node with its children, so that a node step tests up to four boxes
// two binary levels down.  32 words per record: lower x of slots 0..3, upper
// x, lower y, upper y, lower z, upper z (float bits), the slots' words, four
// zero words.  A slot is a child of the binary node that is a leaf (or
// empty), or a grandchild through a child that is inner; an inner slot's
// word names its record as w >> 3; unused slots are empty leaves with boxes
// {+inf, -inf}.  Leaf words and the leaf runs are the binary layout's.
struct WideBuilder {
    const std::vector<uint32_t> &rec2;
    std::vector<uint32_t> rec4;
    uint32_t depth = 0;  // records on the longest root-to-leaf path, minus one
    explicit WideBuilder(const std::vector<uint32_t> &r) : rec2(r) {}
    struct Slot {
        uint32_t w, lo[3], hi[3];
    };
    Slot child(uint32_t p, int ch) const {
        const uint32_t *r = &rec2[16 * (size_t)p];
        return Slot{r[12 + ch], {r[ch], r[4 + ch], r[8 + ch]}, {r[2 + ch], r[6 + ch], r[10 + ch]}};
    }
    static bool inner(uint32_t w) { return (w & 3u) != prt::LEAF_TAG; }
    // The record of binary inner node word w2, at wide depth d: its word.
    uint32_t node(uint32_t w2, uint32_t d) {
        depth = std::max(depth, d);
        const uint32_t p = w2 >> 3;
        Slot slot[4];
        int n = 0;
        for (int ch = 0; ch < 2; ++ch) {
            const Slot c = child(p, ch);
            if (c.w != ((prt::LEAF_EMPTY << 2) | prt::LEAF_TAG) && inner(c.w)) {
                slot[n++] = child(c.w >> 3, 0);
                slot[n++] = child(c.w >> 3, 1);
            } else {
                slot[n++] = c;
            }
        }
        const uint32_t k = (uint32_t)(rec4.size() / 32);
        rec4.resize(rec4.size() + 32, 0u);
        uint32_t words[4];
        for (int j = 0; j < 4; ++j) {
            if (j >= n) {
                words[j] = (prt::LEAF_EMPTY << 2) | prt::LEAF_TAG;
                continue;
            }
            const uint32_t w = slot[j].w;
            words[j] = (w != ((prt::LEAF_EMPTY << 2) | prt::LEAF_TAG) && inner(w)) ? node(w, d + 1) : w;
        }
        uint32_t *r = &rec4[32 * (size_t)k];
        const uint32_t pinf = 0x7F800000u, ninf = 0xFF800000u;
        for (int j = 0; j < 4; ++j)
            for (int a = 0; a < 3; ++a) {
                r[8 * a + j] = j < n ? slot[j].lo[a] : pinf;
                r[8 * a + 4 + j] = j < n ? slot[j].hi[a] : ninf;
            }
        for (int j = 0; j < 4; ++j) r[24 + j] = words[j];
        return k << 3;
    }
};

// The 4-wide records of a binary boxed layout, their root word and the stack
// entries a lane may need (each step stacks at most three slots).
int make_wide_layout(pray_ctx *c, const BoxedBuilder &B, std::vector<uint32_t> &rec4, uint32_t &root,
                     uint32_t &stack_depth) {
    WideBuilder W(B.rec);
    const uint32_t rw = B.rec[13];  // the binary layout's root
    root = (rw != ((prt::LEAF_EMPTY << 2) | prt::LEAF_TAG) && WideBuilder::inner(rw)) ? W.node(rw, 0) : rw;
    if (W.rec4.empty()) W.rec4.assign(32, 0u);
    stack_depth = 3 * (W.depth + 1);
    if (stack_depth > MAX_BOXED_DEPTH)
        return fail(c, PRAY_ERR_UNSUPPORTED, "4-wide boxed layout needs %u stack entries", stack_depth);
    rec4.swap(W.rec4);
    return PRAY_OK;
}

}  // namespace

// Uploads the boxed layout of a tree that takes one (else leaves boxes null).
int build_boxes(pray_ctx *c, SceneRec &sc, const uint32_t *nodes, uint32_t n_nodes, const uint32_t *refs,
                const float *A, const float *AB, const float *AC, hipStream_t st) {
    prt::DevScene &d = sc.d;
    d.boxes = nullptr;
    d.boxes4 = nullptr;
    d.brecs = nullptr;
    d.box_depth = d.box4_depth = d.box4_root = 0;
    if (d.n_tris == 0 || n_nodes < 4) return PRAY_OK;
    BoxedBuilder B;
    if (int rc = make_boxed_layout(c, nodes, refs, A, AB, AC, B)) return rc;
    // the 4-wide form too, unless its stack would be too deep (then boxed
    // node steps stay binary)
    std::vector<uint32_t> rec4;
    uint32_t root4 = 0, depth4 = 0;
    const bool wide = make_wide_layout(nullptr, B, rec4, root4, depth4) == PRAY_OK;
    const size_t n_brefs = B.brefs.size();
    void *rp = nullptr, *fp = nullptr, *bp = nullptr, *wp = nullptr;
    int rc = device_alloc(c, sc, &rp, B.rec.size() * 4);
    if (!rc && wide) rc = device_alloc(c, sc, &wp, rec4.size() * 4);
    if (!rc) rc = device_alloc(c, sc, &fp, (n_brefs + 2) * 4);
    if (!rc) rc = device_alloc(c, sc, &bp, (n_brefs + 2) * 48);  // padded for the ILP look-ahead
    if (rc) return rc;
    HIPC(c, hipMemcpyAsync(rp, B.rec.data(), B.rec.size() * 4, hipMemcpyHostToDevice, st));
    if (wide) HIPC(c, hipMemcpyAsync(wp, rec4.data(), rec4.size() * 4, hipMemcpyHostToDevice, st));
    HIPC(c, hipMemsetAsync(fp, 0, (n_brefs + 2) * 4, st));
    HIPC(c, hipMemcpyAsync(fp, B.brefs.data(), n_brefs * 4, hipMemcpyHostToDevice, st));
    HIPC(c, hipMemsetAsync(bp, 0, (n_brefs + 2) * 48, st));
    hipLaunchKernelGGL(k_pack_records, dim3(nblocks(n_brefs, 256)), dim3(256), 0, st, (const uint32_t *)fp,
                       (uint32_t)n_brefs, d.a, d.ab, d.ac, (float4 *)bp);
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(st));  // the host vectors are freed on return
    d.boxes = static_cast<const uint4 *>(rp);
    d.boxes4 = static_cast<const uint4 *>(wp);
    d.brecs = static_cast<const float4 *>(bp);
    d.box_depth = B.depth;
    d.box4_depth = depth4;
    d.box4_root = root4;
    return PRAY_OK;
}

// Structural check of an uploaded tree so that no kernel can read out of
// bounds or loop: every inner node's child pair is even-aligned and inside
// the array, no path is longer than the declared depth (which also rules out
// cycles), every leaf run lies inside leaf_refs and ends with a
// last-flagged entry, and every referenced triangle exists.
int validate_tree(pray_ctx *c, const uint32_t *nodes, uint32_t n_nodes, const uint32_t *refs, uint32_t n_refs,
                  uint32_t n_tris, uint32_t max_depth) {
    if (n_nodes < 2) return fail(c, PRAY_ERR_INVALID, "tree has no root");
    for (uint32_t k = 0; k < n_refs; k++)
        if ((refs[k] & ~prt::REF_LAST) >= n_tris) return fail(c, PRAY_ERR_INVALID, "leaf ref %u out of range", k);
    if (!(refs[n_refs - 1] & prt::REF_LAST)) return fail(c, PRAY_ERR_INVALID, "last leaf run is unterminated");
    std::vector<std::pair<uint32_t, uint32_t>> stack{{1u, 0u}};
    uint64_t visited = 0;
    while (!stack.empty()) {
        const auto [i, d] = stack.back();
        stack.pop_back();
        if (++visited > n_nodes) return fail(c, PRAY_ERR_INVALID, "tree is not a tree");
        if (d > max_depth) return fail(c, PRAY_ERR_INVALID, "tree deeper than the declared max_depth %u", max_depth);
        const uint32_t w = nodes[2 * i + 1];
        if ((w & 3u) == prt::LEAF_TAG) {
            const uint32_t k = w >> 2;
            if (k == prt::LEAF_EMPTY) continue;
            if (k >= n_refs) return fail(c, PRAY_ERR_INVALID, "leaf of node %u starts outside leaf_refs", i);
        } else {
            const uint32_t lc = w >> 2;
            if ((lc & 1u) || lc < 2 || (uint64_t)lc + 1 >= n_nodes)
                return fail(c, PRAY_ERR_INVALID, "node %u has a bad child index %u", i, lc);
            stack.push_back({lc, d + 1});
            stack.push_back({lc + 1, d + 1});
        }
    }
    return PRAY_OK;
}

int prepare_trace(pray_ctx *c, const SceneRec &sc, uint64_t n) {
    TraceKernel k = pick_kernel(false, sc.tune.layout, sc.tune.stack_lds, sc.tune.leaf_ilp, node_form(sc, sc.tune),
                                sc.tune.leaf_batch > 0);
    return ensure_spill(c, (size_t)launch_shape(c, k, sc.tune, n).lanes * trace_depth(sc, sc.tune) * sizeof(uint4));
}

void tuning_values(const Tuning &t, int *v) {
    v[0] = t.layout, v[1] = t.stack_lds, v[2] = t.leaf_ilp, v[3] = t.fetch_min, v[4] = t.occupancy;
    v[5] = t.wg_waves, v[6] = t.leaf_batch, v[7] = t.node_levels;
}

}  // namespace prc

using namespace prc;

// ------------------------------------------------------------------ C ABI --

extern "C" {

int pray_init(int device, pray_ctx **out) {
    if (!out) return PRAY_ERR_INVALID;
    *out = nullptr;
    // test hook: the status a failed device setup would give
    if (const char *f = std::getenv("PRAY_TEST_INIT_STATUS")) {
        const int code = std::atoi(f);
        if (code < 0) return code;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return PRAY_ERR_NODEVICE;
    if (device < 0 || device >= n) return PRAY_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return PRAY_ERR_HIP;
    auto *c = new pray_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void **)&c->work, (size_t)prt::NCTR * prt::CTR_STRIDE * sizeof(uint32_t)) != hipSuccess) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return PRAY_ERR_HIP;
    }
    c->num_cus = prop.multiProcessorCount;
    if (scratch_setup(device, &c->scratch) != hipSuccess) {
        (void)hipFree(c->work);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return PRAY_ERR_HIP;
    }
    *out = c;
    return PRAY_OK;
}

void pray_destroy(pray_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto &kv : c->scenes) free_scene_rec(*kv.second);
    c->scenes.clear();
    for (auto &kv : c->caches) free_cache_rec(kv.second);
    c->caches.clear();
    for (hipEvent_t &e : c->events)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t &e : c->ievents)
        if (e) (void)hipEventDestroy(e);
    if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
    if (c->join_ev) (void)hipEventDestroy(c->join_ev);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->work) (void)hipFree(c->work);
    for (hipEvent_t e : c->kevents)
        if (e) (void)hipEventDestroy(e);
    if (c->spill) (void)hipFree(c->spill);
    for (void *p : c->ws_ptr)
        if (p) (void)hipFree(p);
    c->build_scratch.reset();
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *pray_last_error(const pray_ctx *c) { return c ? c->err.c_str() : "no context"; }

// Free device memory of the context's device, in bytes (the test hook
// PRAY_TEST_FREE_BYTES replaces the device's figure).
static int free_device_bytes(pray_ctx *c, uint64_t *bytes) {
    size_t fr = 0, tot = 0;
    HIPC(c, hipMemGetInfo(&fr, &tot));
    if (const char *f = std::getenv("PRAY_TEST_FREE_BYTES")) fr = (size_t)std::strtoull(f, nullptr, 10);
    *bytes = fr;
    return PRAY_OK;
}

int pray_mem_ok(pray_ctx *c, uint64_t n_tris, uint64_t n_refs, uint64_t n_nodes, uint64_t n_rays, int *ok) {
    if (!c || !ok) return fail(c, PRAY_ERR_INVALID, "null argument");
    HIPC(c, hipSetDevice(c->device));
    uint64_t fr = 0;
    if (int rc = free_device_bytes(c, &fr)) return rc;
    const uint64_t need = n_nodes * 8 + n_refs * (4 + 48) + n_tris * (36 + 24 + 1) + n_rays * (24 + 8 + 3 + 12 * 4);
    *ok = fr >= need ? 1 : 0;
    return PRAY_OK;
}

int pray_upload_scene(pray_ctx *c, const uint32_t *nodes, uint32_t n_nodes, const uint32_t *leaf_refs,
                      uint32_t n_refs, const float *tri_a, const float *tri_ab, const float *tri_ac,
                      const float *tri_normal, const float *tri_color, const uint8_t *tri_emit, uint32_t n_tris,
                      const float *root_box, uint32_t max_depth, int *scene_id) {
    if (!c || !scene_id) return fail(c, PRAY_ERR_INVALID, "null argument");
    *scene_id = -1;
    if (n_tris > 0 && (!nodes || n_nodes < 2 || !leaf_refs || n_refs == 0 || !tri_a || !tri_ab || !tri_ac || !root_box))
        return fail(c, PRAY_ERR_INVALID, "incomplete scene description");
    if (max_depth > MAX_TREE_DEPTH)
        return fail(c, PRAY_ERR_UNSUPPORTED, "tree depth %u exceeds %u", max_depth, MAX_TREE_DEPTH);
    if (n_tris > 0) {
        const int v = validate_tree(c, nodes, n_nodes, leaf_refs, n_refs, n_tris, max_depth);
        if (v) return v;
    }
    HIPC(c, hipSetDevice(c->device));
    auto sc = std::make_unique<SceneRec>();
    DevScene &d = sc->d;
    d.n_tris = n_tris;
    d.n_nodes = n_nodes;
    d.n_refs = n_refs;
    d.max_depth = max_depth;
    if (root_box) std::memcpy(d.box, root_box, sizeof d.box);
    int rc = PRAY_OK;
    auto up = [&](const void *src, size_t bytes, size_t alloc_bytes, void **dst) -> int {
        *dst = nullptr;
        if (!src || bytes == 0) return PRAY_OK;
        int r = device_alloc(c, *sc, dst, alloc_bytes);
        if (r) return r;
        HIPC(c, hipMemsetAsync(*dst, 0, alloc_bytes, c->stream));
        HIPC(c, hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, c->stream));
        return PRAY_OK;
    };
    void *p = nullptr;
    if (n_tris > 0) {
        // padded to whole 128-byte blocks of 16 nodes
        if ((rc = up(nodes, (size_t)n_nodes * 8, (((size_t)n_nodes + 15) & ~(size_t)15) * 8, &p))) goto bad;
        d.nodes = (const uint2 *)p;
        if ((rc = up(leaf_refs, (size_t)n_refs * 4, (size_t)(n_refs + 2) * 4, &p))) goto bad;
        d.refs = (const uint32_t *)p;
        if ((rc = up(tri_a, (size_t)n_tris * 12, (size_t)n_tris * 12, &p))) goto bad;
        d.a = (const float *)p;
        if ((rc = up(tri_ab, (size_t)n_tris * 12, (size_t)n_tris * 12, &p))) goto bad;
        d.ab = (const float *)p;
        if ((rc = up(tri_ac, (size_t)n_tris * 12, (size_t)n_tris * 12, &p))) goto bad;
        d.ac = (const float *)p;
        if ((rc = up(tri_normal, (size_t)n_tris * 12, (size_t)n_tris * 12, &p))) goto bad;
        d.normal = (const float *)p;
        if ((rc = up(tri_color, (size_t)n_tris * 12, (size_t)n_tris * 12, &p))) goto bad;
        d.color = (const float *)p;
        if ((rc = up(tri_emit, (size_t)n_tris, (size_t)n_tris, &p))) goto bad;
        d.emit = (const uint8_t *)p;
        if ((rc = device_alloc(c, *sc, &p, (size_t)(n_refs + 2) * 48))) goto bad;
        if (hipMemsetAsync(p, 0, (size_t)(n_refs + 2) * 48, c->stream) != hipSuccess) {
            rc = fail(c, PRAY_ERR_HIP, "record clear failed");
            goto bad;
        }
        d.recs = (const float4 *)p;
        hipLaunchKernelGGL(k_pack_records, dim3(nblocks(n_refs, 256)), dim3(256), 0, c->stream, d.refs, n_refs, d.a,
                           d.ab, d.ac, (float4 *)p);
        if (hipGetLastError() != hipSuccess) {
            rc = fail(c, PRAY_ERR_HIP, "record packing launch failed");
            goto bad;
        }
        if ((rc = build_quads(c, *sc, c->stream))) goto bad;
        if ((rc = build_boxes(c, *sc, nodes, n_nodes, leaf_refs, tri_a, tri_ab, tri_ac, c->stream))) goto bad;
        // boxed node steps by default, 4-wide where that form exists: over
        // median trees far ahead of the clip planes; over SAH trees ahead or
        // within a tenth of the KD records, and far ahead where leaves are
        // large (the reference's 32-reference leaves)
        if (d.boxes) sc->tune.node_levels = d.boxes4 ? 4 : 3;
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        rc = fail(c, PRAY_ERR_HIP, "scene upload failed");
        goto bad;
    }
    *scene_id = c->next_id++;
    c->scenes[*scene_id] = std::move(sc);
    return PRAY_OK;
bad:
    (void)hipStreamSynchronize(c->stream);
    free_scene_rec(*sc);
    return rc;
}

int pray_free_scene(pray_ctx *c, int scene_id) {
    if (!c) return PRAY_ERR_INVALID;
    auto it = c->scenes.find(scene_id);
    if (it == c->scenes.end()) return fail(c, PRAY_ERR_NOSCENE, "unknown scene %d", scene_id);
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, hipStreamSynchronize(c->stream));
    free_scene_rec(*it->second);
    c->scenes.erase(it);
    return PRAY_OK;
}

int pray_boxed_layout(const uint32_t *nodes, uint32_t n_nodes, const uint32_t *leaf_refs, uint32_t n_refs,
                      const float *tri_a, const float *tri_ab, const float *tri_ac, uint32_t n_tris,
                      uint32_t max_depth, uint32_t width, uint32_t *n_records, uint32_t *n_brefs, uint32_t *depth,
                      uint32_t *root, uint32_t *rec, uint32_t *brefs) {
    if (!nodes || !leaf_refs || !tri_a || !tri_ab || !tri_ac || !n_records || !n_brefs || !depth || !root ||
        n_tris == 0 || n_refs == 0 || (width != 2 && width != 4))
        return PRAY_ERR_INVALID;
    if (max_depth > MAX_TREE_DEPTH) return PRAY_ERR_UNSUPPORTED;
    if (int v = validate_tree(nullptr, nodes, n_nodes, leaf_refs, n_refs, n_tris, max_depth)) return v;
    if (n_nodes < 4) return PRAY_ERR_UNSUPPORTED;
    BoxedBuilder B;
    if (int rc = make_boxed_layout(nullptr, nodes, leaf_refs, tri_a, tri_ab, tri_ac, B)) return rc;
    *n_brefs = (uint32_t)B.brefs.size();
    if (brefs) std::memcpy(brefs, B.brefs.data(), B.brefs.size() * 4);
    if (width == 2) {
        *n_records = (uint32_t)(B.rec.size() / 16);
        *depth = B.depth;
        *root = B.rec[13];
        if (rec) std::memcpy(rec, B.rec.data(), B.rec.size() * 4);
        return PRAY_OK;
    }
    std::vector<uint32_t> rec4;
    if (int rc = make_wide_layout(nullptr, B, rec4, *root, *depth)) return rc;
    *n_records = (uint32_t)(rec4.size() / 32);
    if (rec) std::memcpy(rec, rec4.data(), rec4.size() * 4);
    return PRAY_OK;
}

int pray_set_tuning(pray_ctx *c, int scene_id, const int *v) {
    if (!c) return PRAY_ERR_INVALID;
    SceneRec *sc = find_scene(c, scene_id);
    if (!sc) return fail(c, PRAY_ERR_NOSCENE, "unknown scene %d", scene_id);
    if (!v) return fail(c, PRAY_ERR_INVALID, "null tuning");
    const int layout = v[0], stack_lds = v[1], leaf_ilp = v[2], fetch_min = v[3], occupancy = v[4], wg_waves = v[5],
              leaf_batch = v[6], node_levels = v[7];
    if (layout >= 0 && layout != PRAY_LAYOUT_INDEXED && layout != PRAY_LAYOUT_PACKED)
        return fail(c, PRAY_ERR_INVALID, "unknown layout %d", layout);
    if (stack_lds >= 0 && stack_lds != 4 && stack_lds != 5 && stack_lds != 6 && stack_lds != 8 && stack_lds != 16)
        return fail(c, PRAY_ERR_INVALID, "stack_lds must be 4, 5, 6, 8 or 16");
    if (leaf_ilp >= 0 && leaf_ilp != 1 && leaf_ilp != 2) return fail(c, PRAY_ERR_INVALID, "leaf_ilp must be 1 or 2");
    if (fetch_min >= 0 && (fetch_min < 1 || fetch_min > WAVE))
        return fail(c, PRAY_ERR_INVALID, "fetch_min must be in 1..64");
    if (occupancy >= 0 && (occupancy < 1 || occupancy > 100))
        return fail(c, PRAY_ERR_INVALID, "occupancy must be in 1..100");
    if (wg_waves >= 0 && wg_waves != 1 && wg_waves != 2 && wg_waves != 4)
        return fail(c, PRAY_ERR_INVALID, "wg_waves must be 1, 2 or 4");
    if (leaf_batch > WAVE) return fail(c, PRAY_ERR_INVALID, "leaf_batch must be in 0..64");
    if ((leaf_batch >= 0 ? leaf_batch : sc->tune.leaf_batch) > 0 &&
        (layout >= 0 ? layout : sc->tune.layout) != PRAY_LAYOUT_PACKED)
        return fail(c, PRAY_ERR_INVALID, "leaf batching traces the packed layout");
    if (node_levels >= 0 && (node_levels < 1 || node_levels > 4))
        return fail(c, PRAY_ERR_INVALID, "node_levels must be 1, 2, 3 or 4");
    const int nl = node_levels >= 0 ? node_levels : sc->tune.node_levels;
    if (nl == 2 && !sc->d.quads)
        return fail(c, PRAY_ERR_UNSUPPORTED, "scene %d has no two-level node records (not a SAH-form tree)", scene_id);
    if (nl == 3 && !sc->d.boxes)
        return fail(c, PRAY_ERR_UNSUPPORTED, "scene %d has no boxed children pairs (its children pairs have equal clips)",
                    scene_id);
    if (nl == 4 && !sc->d.boxes4)
        return fail(c, PRAY_ERR_UNSUPPORTED, "scene %d has no 4-wide boxed layout", scene_id);
    if (nl != 1 && (layout >= 0 ? layout : sc->tune.layout) != PRAY_LAYOUT_PACKED)
        return fail(c, PRAY_ERR_INVALID, "two-level and boxed node steps trace the packed layout");
    Tuning &t = sc->tune;
    if (layout >= 0) t.layout = layout;
    if (stack_lds >= 0) t.stack_lds = stack_lds;
    if (leaf_ilp >= 0) t.leaf_ilp = leaf_ilp;
    if (fetch_min >= 0) t.fetch_min = fetch_min;
    if (occupancy >= 0) t.occupancy = occupancy;
    if (wg_waves >= 0) t.wg_waves = wg_waves;
    if (leaf_batch >= 0) t.leaf_batch = leaf_batch;
    if (node_levels >= 0) t.node_levels = node_levels;
    return PRAY_OK;
}

int pray_get_tuning(pray_ctx *c, int scene_id, int *v) {
    if (!c) return PRAY_ERR_INVALID;
    SceneRec *sc = find_scene(c, scene_id);
    if (!sc) return fail(c, PRAY_ERR_NOSCENE, "unknown scene %d", scene_id);
    if (!v) return fail(c, PRAY_ERR_INVALID, "null tuning");
    tuning_values(sc->tune, v);
    return PRAY_OK;
}

int pray_trace_device(pray_ctx *c, int scene_id, const float *d_rays, uint64_t n, int mode, float det_eps,
                      const float *d_limit2, int32_t *d_hit_idx, float *d_hit_t, void *stream) {
    if (!c) return PRAY_ERR_INVALID;
    SceneRec *sc = find_scene(c, scene_id);
    if (!sc) return fail(c, PRAY_ERR_NOSCENE, "unknown scene %d", scene_id);
    if (n && (!d_rays || !d_hit_idx)) return fail(c, PRAY_ERR_INVALID, "null buffer");
    HIPC(c, hipSetDevice(c->device));
    return enqueue_trace(c, *sc, d_rays, n, nullptr, mode, det_eps, d_limit2, d_hit_idx, d_hit_t, nullptr,
                         pick(c, stream));
}

int pray_trace_counts_device(pray_ctx *c, int scene_id, const float *d_rays, uint64_t n, int mode, float det_eps,
                             const float *d_limit2, int32_t *d_hit_idx, float *d_hit_t, uint32_t *d_counts,
                             void *stream) {
    if (!c) return PRAY_ERR_INVALID;
    SceneRec *sc = find_scene(c, scene_id);
    if (!sc) return fail(c, PRAY_ERR_NOSCENE, "unknown scene %d", scene_id);
    if (n && (!d_rays || !d_hit_idx || !d_counts)) return fail(c, PRAY_ERR_INVALID, "null buffer");
    HIPC(c, hipSetDevice(c->device));
    return enqueue_trace(c, *sc, d_rays, n, nullptr, mode, det_eps, d_limit2, d_hit_idx, d_hit_t, d_counts,
                         pick(c, stream));
}

int pray_trace_device_noalloc(pray_ctx *c, int scene_id, const float *d_rays, uint64_t n, int mode, float det_eps,
                              const float *d_limit2, int32_t *d_hit_idx, float *d_hit_t, void *stream) {
    if (!c) return PRAY_ERR_INVALID;
    SceneRec *sc = find_scene(c, scene_id);
    if (!sc) return fail(c, PRAY_ERR_NOSCENE, "unknown scene %d", scene_id);
    if (n && (!d_rays || !d_hit_idx)) return fail(c, PRAY_ERR_INVALID, "null buffer");
    if (!c->scratch) return fail(c, PRAY_ERR_HIP, "no static scratch on device %d", c->device);
    HIPC(c, hipSetDevice(c->device));
    // The device address looked up at pray_init, never the name `scratch`.
    FixedScratch fx;
    fx.work = static_cast<uint32_t *>(c->scratch);
    fx.spill = reinterpret_cast<uint4 *>(static_cast<char *>(c->scratch) + NOALLOC_COUNTER_BYTES);
    fx.spill_bytes = NOALLOC_SCRATCH_BYTES - NOALLOC_COUNTER_BYTES;
    return enqueue_trace_ex(c, *sc, d_rays, n, nullptr, mode, det_eps, d_limit2, d_hit_idx, d_hit_t, nullptr,
                            pick(c, stream), &fx);
}

int pray_trace(pray_ctx *c, int scene_id, const float *rays, uint64_t n, int mode, float det_eps,
               const float *limit2, int32_t *hit_idx, float *hit_t) {
    if (!c) return PRAY_ERR_INVALID;
    SceneRec *sc = find_scene(c, scene_id);
    if (!sc) return fail(c, PRAY_ERR_NOSCENE, "unknown scene %d", scene_id);
    if (n == 0) return PRAY_OK;
    if (!rays || !hit_idx) return fail(c, PRAY_ERR_INVALID, "null buffer");
    if (mode != PRAY_TRACE_CLOSEST && !limit2) return fail(c, PRAY_ERR_INVALID, "any-hit modes need limit2");
    HIPC(c, hipSetDevice(c->device));
    DBuf dr, dl, di, dt;
    HIPC(c, hipMalloc(&dr.p, n * 24));
    HIPC(c, hipMalloc(&di.p, n * 4));
    HIPC(c, hipMalloc(&dt.p, n * 4));
    HIPC(c, hipMemcpyAsync(dr.p, rays, n * 24, hipMemcpyHostToDevice, c->stream));
    if (mode != PRAY_TRACE_CLOSEST) {
        HIPC(c, hipMalloc(&dl.p, n * 4));
        HIPC(c, hipMemcpyAsync(dl.p, limit2, n * 4, hipMemcpyHostToDevice, c->stream));
    }
    int rc = enqueue_trace(c, *sc, (const float *)dr.p, n, nullptr, mode, det_eps, (const float *)dl.p,
                           (int32_t *)di.p, (float *)dt.p, nullptr, c->stream);
    if (rc) return rc;
    HIPC(c, hipMemcpyAsync(hit_idx, di.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    if (hit_t) HIPC(c, hipMemcpyAsync(hit_t, dt.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return PRAY_OK;
}

int pray_gen_rays_device(pray_ctx *c, uint64_t seed, const float *box6, uint64_t n, float *d_rays, void *stream) {
    if (!c || !box6) return fail(c, PRAY_ERR_INVALID, "null argument");
    if (n > 0xFFFFFFFFull) return fail(c, PRAY_ERR_INVALID, "at most 2^32-1 rays");
    if (n == 0) return PRAY_OK;
    HIPC(c, hipSetDevice(c->device));
    hipLaunchKernelGGL(k_gen_rays, dim3(nblocks(n, 256)), dim3(256), 0, pick(c, stream), (uint32_t)seed,
                       (uint32_t)(seed >> 32), box6[0], box6[1], box6[2], box6[3], box6[4], box6[5], (uint32_t)n,
                       d_rays);
    HIPC(c, hipGetLastError());
    return PRAY_OK;
}

int pray_gen_rays(pray_ctx *c, uint64_t seed, const float *box6, uint64_t n, float *rays) {
    if (!c || !box6 || (n && !rays)) return fail(c, PRAY_ERR_INVALID, "null argument");
    if (n == 0) return PRAY_OK;
    HIPC(c, hipSetDevice(c->device));
    DBuf dr;
    HIPC(c, hipMalloc(&dr.p, n * 24));
    int rc = pray_gen_rays_device(c, seed, box6, n, (float *)dr.p, c->stream);
    if (rc) return rc;
    HIPC(c, hipMemcpyAsync(rays, dr.p, n * 24, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return PRAY_OK;
}

int pray_camera_rays(pray_ctx *c, const float *cam16, uint32_t width, uint32_t height, float *rays) {
    if (!c || !cam16 || !rays) return fail(c, PRAY_ERR_INVALID, "null argument");
    const uint64_t n = (uint64_t)width * height;
    if (n == 0) return PRAY_OK;
    if (n > 0xFFFFFFFFull) return fail(c, PRAY_ERR_INVALID, "image too large");
    HIPC(c, hipSetDevice(c->device));
    Cam cam;
    std::memcpy(&cam, cam16, sizeof cam);
    DBuf dr;
    HIPC(c, hipMalloc(&dr.p, n * 24));
    hipLaunchKernelGGL(prr::k_camera, dim3(nblocks(n, 256)), dim3(256), 0, c->stream, cam, width, height,
                       (float *)dr.p);
    HIPC(c, hipGetLastError());
    HIPC(c, hipMemcpyAsync(rays, dr.p, n * 24, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return PRAY_OK;
}

int pray_render_rt(pray_ctx *c, int scene_id, const float *cam16, uint32_t w, uint32_t h, const float *lights,
                   uint32_t nl, const float *bg, float det_eps, uint8_t *rgb, int32_t *hit_idx) {
    if (!c) return PRAY_ERR_INVALID;
    SceneRec *sc = find_scene(c, scene_id);
    if (!sc) return fail(c, PRAY_ERR_NOSCENE, "unknown scene %d", scene_id);
    if (!cam16 || !bg || !rgb || (nl && !lights)) return fail(c, PRAY_ERR_INVALID, "null argument");
    const uint64_t npix = (uint64_t)w * h;
    if (npix == 0) return PRAY_OK;
    if (npix > (1u << 30) || (uint64_t)nl * npix > (1ull << 31))
        return fail(c, PRAY_ERR_INVALID, "image or light count too large");
    if (sc->d.n_tris && (!sc->d.normal || !sc->d.color))
        return fail(c, PRAY_ERR_INVALID, "scene was uploaded without shading data");
    HIPC(c, hipSetDevice(c->device));
    Range whole("pray_render_rt");
    prr::RtGrid g{};
    g.w = w, g.h = h;
    if (w >= 4 && h >= 2) {
        g.tx = (w - 4) / 4 + 1, g.ty = (h - 2) / 2 + 1, g.missing = w % 4, g.rows_sc = h - 1;
        g.n_avx = 8 * g.tx * g.ty;
        g.n_sc = g.rows_sc * (g.missing + 1) + ((h & 1u) ? w : 0u);
    } else {  // too small for a packet: every pixel takes the scalar path
        g.tx = g.ty = 0, g.missing = w - 1, g.rows_sc = h, g.n_avx = 0, g.n_sc = w * h;
    }
    const uint64_t lanes = (uint64_t)g.n_avx + g.n_sc;
    const uint64_t nsh = lanes * nl;
    Cam cam;
    std::memcpy(&cam, cam16, sizeof cam);
    // Buffers come from the context's workspace (slots 64..71), which keeps
    // them across frames: a frame allocates only when it outgrows them.
    struct {
        void *p = nullptr;
    } rays, hit, tt, lt, sr, sl, so, out;
    WS_GET(c, 64, rays.p, npix * 24);
    WS_GET(c, 65, hit.p, npix * 4);
    WS_GET(c, 66, tt.p, npix * 4);
    WS_GET(c, 67, out.p, npix * 3);
    if (nl) {
        WS_GET(c, 68, lt.p, (size_t)nl * 24);
        WS_GET(c, 69, sr.p, nsh * 24);
        WS_GET(c, 70, sl.p, nsh * 4);
        WS_GET(c, 71, so.p, nsh * 4);
        HIPC(c, hipMemcpyAsync(lt.p, lights, (size_t)nl * 24, hipMemcpyHostToDevice, c->stream));
    }
    {
        Range r("rt generate");
        hipLaunchKernelGGL(prr::k_camera, dim3(nblocks(npix, 256)), dim3(256), 0, c->stream, cam, w, h,
                           (float *)rays.p);
        HIPC(c, hipGetLastError());
    }
    int rc;
    {
        Range r("rt extend");
        rc = enqueue_trace(c, *sc, (const float *)rays.p, npix, nullptr, PRAY_TRACE_CLOSEST, det_eps, nullptr,
                           (int32_t *)hit.p, (float *)tt.p, nullptr, c->stream);
    }
    if (rc) return rc;
    prr::RtArgs a;
    a.g = g;
    a.rays = (const float *)rays.p;
    a.hit = (const int32_t *)hit.p;
    a.t = (const float *)tt.p;
    a.lights = (const float *)lt.p;
    a.nl = nl;
    a.bg[0] = bg[0], a.bg[1] = bg[1], a.bg[2] = bg[2];
    a.srays = (float *)sr.p;
    a.slim = (float *)sl.p;
    a.socc = (const int32_t *)so.p;
    a.rgb = (uint8_t *)out.p;
    if (nl && sc->d.n_tris) {
        Range r("rt shadow");
        hipLaunchKernelGGL(prr::k_rt_shadow_setup, dim3(nblocks(lanes, 256)), dim3(256), 0, c->stream, sc->d, a);
        HIPC(c, hipGetLastError());
        rc = enqueue_trace(c, *sc, (const float *)sr.p, nsh, nullptr, PRAY_TRACE_ANY_INCLUSIVE, det_eps,
                           (const float *)sl.p, (int32_t *)so.p, nullptr, nullptr, c->stream);
        if (rc) return rc;
    } else if (nl) {
        hipLaunchKernelGGL(k_fill_miss, dim3(nblocks(nsh, 256)), dim3(256), 0, c->stream, (uint32_t)nsh,
                           (int32_t *)so.p, (float *)nullptr, (uint32_t *)nullptr);
    }
    {
        Range r("rt shade");
        hipLaunchKernelGGL(prr::k_rt_shade, dim3(nblocks(lanes, WAVE)), dim3(WAVE), 0, c->stream, sc->d, a);
        HIPC(c, hipGetLastError());
    }
    HIPC(c, hipMemcpyAsync(rgb, out.p, npix * 3, hipMemcpyDeviceToHost, c->stream));
    if (hit_idx) HIPC(c, hipMemcpyAsync(hit_idx, hit.p, npix * 4, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return PRAY_OK;
}

int pray_render_pt(pray_ctx *c, int scene_id, const float *cam16, uint32_t w, uint32_t h, uint32_t spp,
                   uint32_t max_depth, uint32_t rr_depth, uint64_t seed, const float *bg, float det_eps, float *hdr,
                   uint8_t *rgb) {
    return pray_render_pt_ex(c, scene_id, cam16, w, h, spp, max_depth, rr_depth, seed, bg, det_eps, hdr, rgb, 1);
}

namespace {

// Samples first .. first+spp-1 of a path-traced image.  The running sum of
// per-sample radiance (3 floats per pixel) starts from accum_in (zeros when
// first == 0) and is returned in accum_out; hdr / rgb receive the mean over
// all first+spp samples.  Samples are added in order into the same fp32 sum,
// so splitting a render into several calls gives the image of one call.
int render_pt(pray_ctx *c, int scene_id, const float *cam16, uint32_t w, uint32_t h, uint32_t first, uint32_t spp,
              uint32_t max_depth, uint32_t rr_depth, uint64_t seed, const float *bg, float det_eps,
              const float *accum_in, float *accum_out, float *hdr, uint8_t *rgb, int use_graph) {
    if (!c) return PRAY_ERR_INVALID;
    SceneRec *sc = find_scene(c, scene_id);
    if (!sc) return fail(c, PRAY_ERR_NOSCENE, "unknown scene %d", scene_id);
    if (!cam16 || !bg) return fail(c, PRAY_ERR_INVALID, "null argument");
    const uint64_t npix = (uint64_t)w * h;
    if (npix == 0 || spp == 0) return PRAY_OK;
    if (npix > (1u << 28)) return fail(c, PRAY_ERR_INVALID, "image too large");
    if (first && !accum_in) return fail(c, PRAY_ERR_INVALID, "resuming at sample %u needs the running sum", first);
    // The mean divides by the sample count as a float: keep it exact.
    if ((uint64_t)first + spp > (1u << 24)) return fail(c, PRAY_ERR_INVALID, "more than 2^24 samples");
    if (sc->d.n_tris && (!sc->d.normal || !sc->d.color || !sc->d.emit))
        return fail(c, PRAY_ERR_INVALID, "scene was uploaded without shading data");
    HIPC(c, hipSetDevice(c->device));
    Cam cam;
    std::memcpy(&cam, cam16, sizeof cam);
    if (max_depth > 64) return fail(c, PRAY_ERR_INVALID, "max_depth must be at most 64");
    Range whole("pray_render_pt");
    // Samples per pass: as many as keep the paths in flight within the
    // context's path budget (at least one), so that every bounce's trace
    // launch holds samples x pixels paths.  Path p of a pass is sample
    // p / npix of pixel p % npix; results do not depend on the split.
    // The path buffers (100 bytes per path: two queues of rays and states,
    // hit index and distance, contribution) also stay within half the free
    // device memory, counting what the workspace already holds for them.
    uint64_t per_pass = std::max<uint64_t>(1, std::min<uint64_t>(c->path_budget, 1ull << 31) / npix);
    {
        uint64_t fr = 0;
        if (int rc = free_device_bytes(c, &fr)) return rc;
        uint64_t held = 0;
        for (size_t k = 1; k <= 9 && k < c->ws_bytes.size(); ++k) held += c->ws_bytes[k];
        per_pass = std::min<uint64_t>(per_pass, std::max<uint64_t>(1, (fr / 2 + held) / 100 / npix));
    }
    const uint32_t S = (uint32_t)std::min<uint64_t>(spp, per_pass);
    const uint64_t npath = npix * S;
    // Buffers from the context's render workspace (slots 0..14), reused by
    // later calls.
    struct Buf {
        void *p = nullptr;
    } prim, rq[2], sq[2], cnt, hit, tt, contrib, sum, outh, outb, qsum, phit, pt;
    WS_GET(c, 0, prim.p, npix * 24);
    WS_GET(c, 13, phit.p, npix * 4);  // the camera rays' hits, shared by all samples
    WS_GET(c, 14, pt.p, npix * 4);
    for (int q = 0; q < 2; q++) {
        WS_GET(c, 1 + q, rq[q].p, npath * 24);
        WS_GET(c, 3 + q, sq[q].p, npath * 16);
    }
    WS_GET(c, 5, cnt.p, 6 * sizeof(uint32_t));    // count0, count1, sample, pad, traced rays (u64)
    WS_GET(c, 6, qsum.p, 65 * sizeof(uint64_t));  // paths traced per bounce, all samples
    WS_GET(c, 7, hit.p, npath * 4);
    WS_GET(c, 8, tt.p, npath * 4);
    WS_GET(c, 9, contrib.p, npath * 12);
    WS_GET(c, 10, sum.p, npix * 12);
    WS_GET(c, 11, outh.p, npix * 12);
    if (rgb) WS_GET(c, 12, outb.p, npix * 3);
    uint32_t *count = (uint32_t *)cnt.p;
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(prr::k_camera, dim3(nblocks(npix, 256)), dim3(256), 0, st, cam, w, h, (float *)prim.p);
    if (first)
        HIPC(c, hipMemcpyAsync(sum.p, accum_in, npix * 12, hipMemcpyHostToDevice, st));
    else
        hipLaunchKernelGGL(prr::k_fill_f32, dim3(nblocks(3 * npix, 256)), dim3(256), 0, st, (float *)sum.p,
                           (uint32_t)(3 * npix), 0.f);
    hipLaunchKernelGGL(prt::k_zero_u32, dim3(1), dim3(64), 0, st, count, 6u);
    hipLaunchKernelGGL(prt::k_zero_u32, dim3(3), dim3(64), 0, st, (uint32_t *)qsum.p, 130u);
    hipLaunchKernelGGL(prt::k_set_u32, dim3(1), dim3(64), 0, st, count + 2, first);  // sample index
    HIPC(c, hipGetLastError());
    unsigned long long *traced = (unsigned long long *)(count + 4);
    prr::PtArgs P;
    P.n_pix = (uint32_t)npix;
    P.max_depth = max_depth;
    P.rr_depth = rr_depth;
    P.k0 = (uint32_t)seed, P.k1 = (uint32_t)(seed >> 32);
    P.bg[0] = bg[0], P.bg[1] = bg[1], P.bg[2] = bg[2];
    P.sample = count + 2;
    // the reference's estimator choice: diffuse at a depth limit of at most 1
    // (scene_preprocessor.cpp:34-40)
    P.diffuse = max_depth <= 1 ? 1u : 0u;
    // Every sample of a pixel starts with the pixel's camera ray: its hit is
    // traced once for the whole call (and counted once among the rays).
    {
        Range r("pt camera hits");
        if (sc->d.n_tris) {
            int rc = enqueue_trace(c, *sc, (const float *)prim.p, npix, nullptr, PRAY_TRACE_CLOSEST, det_eps, nullptr,
                                   (int32_t *)phit.p, (float *)pt.p, nullptr, st);
            if (rc) return rc;
        } else {
            hipLaunchKernelGGL(k_fill_miss, dim3(nblocks(npix, 256)), dim3(256), 0, st, (uint32_t)npix,
                               (int32_t *)phit.p, (float *)pt.p, (uint32_t *)nullptr);
        }
        hipLaunchKernelGGL(prt::k_set_u32, dim3(1), dim3(64), 0, st, count + 3, (uint32_t)npix);
        hipLaunchKernelGGL(prr::k_acc_rays, dim3(1), dim3(64), 0, st, traced, (const uint32_t *)(count + 3));
    }
    // The persistent trace kernel needs its spill area before any capture.
    {
        int rc = prepare_trace(c, *sc, npath);
        if (rc) return rc;
    }
    // One pass of `ns` samples: generate, (zero next count, extend, shade)
    // per bounce (the first bounce's hits are the camera rays'), accumulate.
    auto pass = [&](uint32_t ns) -> int {
        const uint64_t np = npix * ns;
        P.samples = ns;
        Range rp("pt pass");
        {
            Range rg("pt generate");
            hipLaunchKernelGGL(prr::k_pt_generate, dim3(nblocks(np, 256)), dim3(256), 0, st, (uint32_t)np,
                               (float4 *)sq[0].p, count + 0);
        }
        for (uint32_t b = 0; b <= max_depth; b++) {
            const int cur = b & 1, nxt = cur ^ 1;
            hipLaunchKernelGGL(prt::k_zero_u32, dim3(1), dim3(64), 0, st, count + nxt, 1u);
            hipLaunchKernelGGL(prr::k_acc_rays, dim3(1), dim3(64), 0, st, (unsigned long long *)qsum.p + b,
                               (const uint32_t *)(count + cur));
            if (b > 0) {
                hipLaunchKernelGGL(prr::k_acc_rays, dim3(1), dim3(64), 0, st, traced, (const uint32_t *)(count + cur));
                if (sc->d.n_tris) {
                    Range r("pt extend");
                    int rc = enqueue_trace(c, *sc, (const float *)rq[cur].p, np, count + cur, PRAY_TRACE_CLOSEST,
                                           det_eps, nullptr, (int32_t *)hit.p, (float *)tt.p, nullptr, st);
                    if (rc) return rc;
                } else {
                    hipLaunchKernelGGL(k_fill_miss, dim3(nblocks(np, 256)), dim3(256), 0, st, (uint32_t)np,
                                       (int32_t *)hit.p, (float *)tt.p, (uint32_t *)nullptr);
                }
            }
            Range r("pt shade");
            const bool first_bounce = b == 0;
            hipLaunchKernelGGL(prr::k_pt_shade<prr::PT_PLAIN>, dim3(nblocks(np, 256)), dim3(256), 0, st, sc->d, P,
                               prr::CacheArgs{}, b, first_bounce ? 1 : 0,
                               (const float *)(first_bounce ? prim.p : rq[cur].p), (const float4 *)sq[cur].p,
                               (const uint32_t *)(count + cur), (const int32_t *)(first_bounce ? phit.p : hit.p),
                               (const float *)(first_bounce ? pt.p : tt.p), (float *)rq[nxt].p, (float4 *)sq[nxt].p,
                               count + nxt, (float *)contrib.p, prr::BatchArgs{});
        }
        Range ra("pt accumulate");
        hipLaunchKernelGGL(prr::k_pt_accumulate, dim3(nblocks(3 * npix, 256)), dim3(256), 0, st, (uint32_t)npix, ns,
                           (const float *)contrib.p, (float *)sum.p, count + 2);
        HIPC(c, hipGetLastError());
        return PRAY_OK;
    };
    const uint32_t full = spp / S, rest = spp % S;
    if (use_graph && c->use_graphs && full > 1) {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        HIPC(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = pass(S);
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
        for (uint32_t s = 0; s < full && e == hipSuccess; s++) e = hipGraphLaunch(exec, st);
        const hipError_t e2 = hipStreamSynchronize(st);
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess || e2 != hipSuccess)
            return fail(c, PRAY_ERR_HIP, "graph launch failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    } else {
        for (uint32_t s = 0; s < full; s++) {
            int rc = pass(S);
            if (rc) return rc;
        }
    }
    if (rest) {  // the last samples, fewer than a pass holds
        int rc = pass(rest);
        if (rc) return rc;
    }
    if (P.diffuse)
        hipLaunchKernelGGL(prr::k_pt_finalize_diffuse, dim3(nblocks(3 * npix, 256)), dim3(256), 0, st,
                           (uint32_t)npix, (const float *)sum.p, (float)(first + spp), (const int32_t *)phit.p,
                           sc->d.emit, sc->d.color, bg[0], bg[1], bg[2], max_depth, (float *)outh.p,
                           (uint8_t *)outb.p);
    else
        hipLaunchKernelGGL(prr::k_pt_finalize, dim3(nblocks(3 * npix, 256)), dim3(256), 0, st, (uint32_t)npix,
                           (const float *)sum.p, (float)(first + spp), (float *)outh.p, (uint8_t *)outb.p);
    HIPC(c, hipGetLastError());
    if (hdr) HIPC(c, hipMemcpyAsync(hdr, outh.p, npix * 12, hipMemcpyDeviceToHost, st));
    if (rgb) HIPC(c, hipMemcpyAsync(rgb, outb.p, npix * 3, hipMemcpyDeviceToHost, st));
    if (accum_out) HIPC(c, hipMemcpyAsync(accum_out, sum.p, npix * 12, hipMemcpyDeviceToHost, st));
    HIPC(c, hipMemcpyAsync(&c->last_rays, traced, 8, hipMemcpyDeviceToHost, st));
    c->last_queues.assign(max_depth + 1, 0);
    HIPC(c, hipMemcpyAsync(c->last_queues.data(), qsum.p, 8 * ((size_t)max_depth + 1), hipMemcpyDeviceToHost, st));
    HIPC(c, hipStreamSynchronize(st));
    return PRAY_OK;
}

}  // namespace

int pray_render_pt_ex(pray_ctx *c, int scene_id, const float *cam16, uint32_t w, uint32_t h, uint32_t spp,
                      uint32_t max_depth, uint32_t rr_depth, uint64_t seed, const float *bg, float det_eps, float *hdr,
                      uint8_t *rgb, int use_graph) {
    return render_pt(c, scene_id, cam16, w, h, 0, spp, max_depth, rr_depth, seed, bg, det_eps, nullptr, nullptr, hdr,
                     rgb, use_graph);
}

int pray_render_pt_progressive(pray_ctx *c, int scene_id, const float *cam16, uint32_t w, uint32_t h,
                               uint32_t first_sample, uint32_t spp, uint32_t max_depth, uint32_t rr_depth,
                               uint64_t seed, const float *bg, float det_eps, float *accum, float *hdr, uint8_t *rgb) {
    if (c && !accum) return fail(c, PRAY_ERR_INVALID, "null accumulation buffer");
    return render_pt(c, scene_id, cam16, w, h, first_sample, spp, max_depth, rr_depth, seed, bg, det_eps, accum,
                     accum, hdr, rgb, 1);
}

int pray_device_alloc(pray_ctx *c, uint64_t bytes, void **p) {
    if (!c || !p) return fail(c, PRAY_ERR_INVALID, "null argument");
    *p = nullptr;
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, hipMalloc(p, bytes ? bytes : 1));
    return PRAY_OK;
}

int pray_device_free(pray_ctx *c, void *p) {
    if (!c) return PRAY_ERR_INVALID;
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, hipStreamSynchronize(c->stream));
    HIPC(c, hipFree(p));
    return PRAY_OK;
}

int pray_host_alloc(pray_ctx *c, uint64_t bytes, void **p) {
    if (!c || !p) return fail(c, PRAY_ERR_INVALID, "null argument");
    *p = nullptr;
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault));
    return PRAY_OK;
}

int pray_host_free(void *p) {
    if (!p) return PRAY_OK;
    return hipHostFree(p) == hipSuccess ? PRAY_OK : PRAY_ERR_HIP;
}

int pray_copy_to_device(pray_ctx *c, void *dst, const void *src, uint64_t bytes) {
    if (!c || (bytes && (!dst || !src))) return fail(c, PRAY_ERR_INVALID, "null argument");
    if (!bytes) return PRAY_OK;
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return PRAY_OK;
}

int pray_copy_to_host(pray_ctx *c, void *dst, const void *src, uint64_t bytes) {
    if (!c || (bytes && (!dst || !src))) return fail(c, PRAY_ERR_INVALID, "null argument");
    if (!bytes) return PRAY_OK;
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return PRAY_OK;
}

int pray_sync(pray_ctx *c) {
    if (!c) return PRAY_ERR_INVALID;
    HIPC(c, hipSetDevice(c->device));
    HIPC(c, hipStreamSynchronize(c->stream));
    return PRAY_OK;
}

int pray_copy_device(pray_ctx *c, void *dst, const void *src, uint64_t bytes, void *stream) {
    if (!c || (bytes && (!dst || !src)) || (bytes & 15u)) return fail(c, PRAY_ERR_INVALID, "bad copy arguments");
    if (!bytes) return PRAY_OK;
    HIPC(c, hipSetDevice(c->device));
    const uint64_t n4 = bytes / 16;
    const unsigned grid = (unsigned)std::min<uint64_t>((n4 + 255) / 256, (uint64_t)c->num_cus * 32);
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, pick(c, stream), (uint4 *)dst, (const uint4 *)src, n4);
    HIPC(c, hipGetLastError());
    return PRAY_OK;
}

int pray_event_record(pray_ctx *c, int slot) {
    if (!c || slot < 0 || slot >= 16) return fail(c, PRAY_ERR_INVALID, "bad event slot");
    HIPC(c, hipSetDevice(c->device));
    if (!c->events[slot]) HIPC(c, hipEventCreate(&c->events[slot]));
    HIPC(c, hipEventRecord(c->events[slot], c->stream));
    return PRAY_OK;
}

int pray_set_path_budget(pray_ctx *c, uint64_t max_paths) {
    if (!c) return PRAY_ERR_INVALID;
    c->path_budget = max_paths;
    return PRAY_OK;
}

int pray_use_graphs(pray_ctx *c, int enable) {
    if (!c) return PRAY_ERR_INVALID;
    c->use_graphs = enable != 0;
    return PRAY_OK;
}

int pray_kernel_timing(pray_ctx *c, int enable) {
    if (!c) return PRAY_ERR_INVALID;
    HIPC(c, hipSetDevice(c->device));
    for (hipEvent_t &e : c->kevents)
        if (!e) HIPC(c, hipEventCreate(&e));
    if (enable && !c->kernel_timing) c->kt_total = 0;
    c->kernel_timing = enable != 0;
    return PRAY_OK;
}

int pray_last_kernel_ms(pray_ctx *c, float *ms) {
    if (!c || !ms) return fail(c, PRAY_ERR_INVALID, "null argument");
    if (!c->kt_total) return fail(c, PRAY_ERR_INVALID, "no traversal kernel has been timed");
    HIPC(c, hipSetDevice(c->device));
    const int k = (int)((c->kt_total - 1) % pray_ctx::KT_RING);
    HIPC(c, hipEventSynchronize(c->kevents[2 * k + 1]));
    HIPC(c, hipEventElapsedTime(ms, c->kevents[2 * k], c->kevents[2 * k + 1]));
    return PRAY_OK;
}

int pray_timed_kernels_ms(pray_ctx *c, float *mean_ms, uint32_t *count) {
    if (!c || !mean_ms || !count) return fail(c, PRAY_ERR_INVALID, "null argument");
    if (!c->kt_total) return fail(c, PRAY_ERR_INVALID, "no traversal kernel has been timed");
    HIPC(c, hipSetDevice(c->device));
    const uint64_t n = std::min<uint64_t>(c->kt_total, pray_ctx::KT_RING);
    HIPC(c, hipEventSynchronize(c->kevents[2 * ((c->kt_total - 1) % pray_ctx::KT_RING) + 1]));
    double sum = 0;
    for (uint64_t i = c->kt_total - n; i < c->kt_total; i++) {
        const int k = (int)(i % pray_ctx::KT_RING);
        float ms = 0;
        HIPC(c, hipEventElapsedTime(&ms, c->kevents[2 * k], c->kevents[2 * k + 1]));
        sum += ms;
    }
    *mean_ms = (float)(sum / (double)n);
    *count = (uint32_t)n;
    return PRAY_OK;
}

int pray_event_elapsed(pray_ctx *c, int s0, int s1, float *ms) {
    if (!c || !ms || s0 < 0 || s0 >= 16 || s1 < 0 || s1 >= 16 || !c->events[s0] || !c->events[s1])
        return fail(c, PRAY_ERR_INVALID, "bad event slot");
    HIPC(c, hipEventSynchronize(c->events[s1]));
    HIPC(c, hipEventElapsedTime(ms, c->events[s0], c->events[s1]));
    return PRAY_OK;
}

}  // extern "C"
