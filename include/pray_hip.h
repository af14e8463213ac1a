/*
 * pray_hip.h -- C ABI of libpray_hip.so, the MI355X (gfx950) tracer of the
 * Nostrum Pray drop-in.
 *
 * The reference crosses to the GPU through C++ entry points that allocate,
 * copy, launch and free per call and terminate the process on any error:
 *   bool isComputableOnGPU(const Scene&)            include/path_trace.cu.h:11
 *   void pathTraceOnGPURecursive(Image&, const Scene&) include/path_trace.cu.h:12
 *   KDTreeGPU buildGPUTreeFromKDTree(...)           include/gpu_util.h:10
 *   class KDTreeGPU / struct KDTreeGPUNode          include/kd_tree_gpu.h:11-19,
 *                                                   include/kd_tree_gpu_structs.h:7-22
 *   SpatialDatabase::intersect (closest hit)        include/spatial.h:15, src/kd_tree.cpp:142-279
 *   shadow test through intersect                   src/pray.cpp:39-44, :123-133
 *   void render(Image&, const Scene&)               include/pray.h:50
 * This library replaces them with plain-C calls: plain pointers and sizes,
 * int status codes (never exit()), device memory owned by an opaque context,
 * scenes uploaded once and reused.
 *
 * Conventions
 *   - Every function returns PRAY_OK (0) or a negative PRAY_ERR_* code; the
 *     text of the last error is available from pray_last_error().
 *   - Host-buffer calls are synchronous: results are on the host when they
 *     return.  *_device calls take device pointers and a hipStream_t (passed
 *     as void*, NULL = the context's stream) and return after enqueueing.
 *   - One context per host thread; a context is not re-entrant.
 *   - Floating-point results are computed in FP32 without contraction and are
 *     bit-identical to the CPU restatement of the reference's algorithm (see
 *     oracle/).  Ray directions are taken as given (callers pass directions
 *     that are already normalised, as the reference's Ray constructor does,
 *     geometry.h:84-87).
 */
#ifndef PRAY_HIP_H
#define PRAY_HIP_H

#include <stdint.h>

#ifndef PRAY_API
#define PRAY_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define PRAY_OK 0
#define PRAY_ERR_INVALID (-1)     /* bad argument */
#define PRAY_ERR_HIP (-2)         /* HIP runtime error */
#define PRAY_ERR_NOMEM (-3)       /* device allocation failed */
#define PRAY_ERR_NOSCENE (-4)     /* unknown scene id */
#define PRAY_ERR_UNSUPPORTED (-5) /* e.g. tree deeper than the traversal stack */
#define PRAY_ERR_NODEVICE (-6)    /* no GPU visible */

/* Trace modes. */
#define PRAY_TRACE_CLOSEST 0        /* lexicographic minimum of (t, triangle index) over t > 0 */
#define PRAY_TRACE_ANY 1            /* occluded iff some hit has t > 0 and t*t <  limit2 (pray.cpp:42) */
#define PRAY_TRACE_ANY_INCLUSIVE 2  /* occluded iff some hit has t > 0 and t*t <= limit2 (pray.cpp:129) */

/* Triangle data layouts of an uploaded scene (tuning; results are identical). */
#define PRAY_LAYOUT_INDEXED 0  /* leaf index arrays + per-field triangle arrays a/ab/ac */
#define PRAY_LAYOUT_PACKED 1   /* per-leaf packed 48-byte triangle records */

typedef struct pray_ctx pray_ctx;

/* Opens `device` and creates a context with its own stream.  Replaces the
 * implicit device setup of path_trace.cu:501-504.
 * Test hook: with the environment variable PRAY_TEST_INIT_STATUS set to a
 * negative PRAY_ERR_* code, pray_init returns that code without touching the
 * device, so callers' handling of a failed setup can be tested anywhere. */
PRAY_API int pray_init(int device, pray_ctx **out);
PRAY_API void pray_destroy(pray_ctx *ctx);
PRAY_API const char *pray_last_error(const pray_ctx *ctx);

/* Memory feasibility for the flat layout (replaces isComputableOnGPU,
 * path_trace.cu:557-583): *ok = 1 when the device has room for the scene
 * (nodes, leaf refs, triangles incl. shading data) and n_rays rays with
 * their hit records and image.  Does not allocate.
 * Test hook: PRAY_TEST_FREE_BYTES in the environment replaces the free
 * device memory the answer is judged by. */
PRAY_API int pray_mem_ok(pray_ctx *ctx, uint64_t n_tris, uint64_t n_refs, uint64_t n_nodes, uint64_t n_rays, int *ok);

/* Uploads a flattened tree once (replaces copyKDTree, path_trace.cu:480-492).
 *   nodes      n_nodes 8-byte nodes as 2 uint32 words each (layout in
 *              nostrum_pray_amd/host/gpu_layout.hpp); node 1 is the root.
 *   leaf_refs  n_refs uint32: triangle index | 0x80000000 on a leaf's last entry
 *   tri_a, tri_ab, tri_ac   3*n_tris floats each: vertex a and edges b-a, c-a
 *   tri_normal, tri_color   3*n_tris floats each, may be NULL for trace-only scenes
 *   tri_emit   n_tris bytes (1 = emissive), may be NULL
 *   root_box   6 floats: min xyz, max xyz of the root cell
 *   max_depth  root-to-leaf depth of the tree (<= 64)
 * n_tris == 0 is a valid empty scene (every ray misses).  The upload also
 * makes the boxed layouts of pray_boxed_layout and selects the 4-wide one
 * (node_levels 4). */
PRAY_API int pray_upload_scene(pray_ctx *ctx, const uint32_t *nodes, uint32_t n_nodes, const uint32_t *leaf_refs,
                      uint32_t n_refs, const float *tri_a, const float *tri_ab, const float *tri_ac,
                      const float *tri_normal, const float *tri_color, const uint8_t *tri_emit, uint32_t n_tris,
                      const float *root_box, uint32_t max_depth, int *scene_id);
PRAY_API int pray_free_scene(pray_ctx *ctx, int scene_id);

/* The boxed layout the upload makes for boxed node steps (no reference
 * counterpart; host code only, no device or context needed).  For a valid
 * tree (the arrays of pray_upload_scene): every inner node becomes a pair
 * record of its two children's boxes, every leaf of more than 2 references
 * is split further (triangles
 * sorted by centroid along the longest axis of their centroids' bounds, ties
 * by index, halved until no leaf holds more), and the leaves' runs go to a
 * reference array of their own (brefs: n_brefs words, the last entry of each
 * run | 0x80000000).  A box is the union of its triangles' boxes, vertices a,
 * a + ab, a + ac summed in double and rounded outward, cut to the node's
 * cell: where a children pair's two clips are equal (a KD split plane), child
 * 0's cell ends at that plane on the parent's axis and child 1's begins
 * there; a pair with two different clips (a median tree's) cuts nothing.
 * {+inf, -inf} when empty.  A leaf word is (first entry of its run in brefs) << 2 | 3;
 * 0xFFFFFFFF is an empty leaf; an inner word is (record index) << 3.
 *   width 2  (node_levels 3) records of 16 uint32: per axis x, y, z {child 0
 *            lower, child 1 lower, child 0 upper, child 1 upper} (float
 *            bits), the two children's words, two zero words.  Record 0's
 *            child 0 is empty and its child 1 is the root (*root receives its
 *            word); *depth receives the root-to-leaf edges.
 *   width 4  (node_levels 4) records of 32 uint32, each a binary node
 *            collapsed with its children: up to four slots (a leaf child, or
 *            the two children of an inner child), {lower x of slots 0..3},
 *            {upper x}, {lower y}, {upper y}, {lower z}, {upper z}, the
 *            slots' words, four zero words; unused slots are empty.  *root
 *            receives the root word and *depth the stack entries a lane may
 *            need (three per record level).
 * *n_records receives the record count; rec (16 or 32 * n_records words) and
 * brefs are filled when not NULL.  PRAY_ERR_UNSUPPORTED for a tree of fewer
 * than two levels or whose layout would be too deep, PRAY_ERR_INVALID for a
 * malformed one. */
PRAY_API int pray_boxed_layout(const uint32_t *nodes, uint32_t n_nodes, const uint32_t *leaf_refs, uint32_t n_refs,
                               const float *tri_a, const float *tri_ab, const float *tri_ac, uint32_t n_tris,
                               uint32_t max_depth, uint32_t width, uint32_t *n_records, uint32_t *n_brefs,
                               uint32_t *depth, uint32_t *root, uint32_t *rec, uint32_t *brefs);

/* Tree build on the GPU (the reference builds its trees on the CPU,
 * kd_tree.cpp:25-84, kd_tree_sah.cpp:32-296), built and kept on the device
 * in the traversal's layout.  Any tree gives the same closest hits.
 *   method     PRAY_BUILD_SAH: binned-SAH KD tree built level by level (64
 *              bins per axis, the reference's SAH cost with its empty-side
 *              factor; references of straddling triangles are duplicated);
 *              PRAY_BUILD_MORTON: bounding interval hierarchy over the
 *              triangles in the order of the 30-bit Morton code of their box
 *              centres (binary radix tree, Karras 2012): the fastest build,
 *              a slower tree
 *   tris       n x 9 floats a, b, c (host memory)
 *   color      n x 3 floats, NULL for a scene without shading data
 *   emit       n bytes, may be NULL (no emitters)
 *   leaf_size  nodes of at most this many triangles are not split (1..64)
 *   max_depth  SAH depth cap (0 = 40, at most 64)
 *   traversal_cost, intersection_cost  SAH costs (<= 0: 2 and 1)
 *   build_ms   optional: device time from the uploaded triangles to the
 *              finished layout
 * The new scene takes the default tuning.  The context keeps the builds'
 * device scratch, grown to the largest build so far, until pray_destroy: a
 * scene rebuilt every frame allocates only its own arrays after the first
 * build.  The largest part is the SAH level histograms, 1,536 bytes per node
 * of the widest tree level. */
#define PRAY_BUILD_MORTON 0
#define PRAY_BUILD_SAH 1
PRAY_API int pray_build_scene(pray_ctx *ctx, const float *tris, const float *color, const uint8_t *emit, uint32_t n,
                              int method, uint32_t leaf_size, uint32_t max_depth, float traversal_cost,
                              float intersection_cost, int *scene_id, float *build_ms);
/* A scene's device layout: sizes[4] = {n_nodes, n_refs, n_tris, max_depth};
 * nodes = n_nodes x 2 words, refs = n_refs words, box = 6 floats (any of the
 * buffers may be NULL). */
PRAY_API int pray_scene_layout(pray_ctx *ctx, int scene_id, uint64_t *sizes, uint32_t *nodes, uint32_t *refs,
                               float *box);

/* Tuning of a scene's traversal (results never depend on it), as an array
 * tuning[PRAY_TUNING_FIELDS]:
 *   [0] layout     PRAY_LAYOUT_*
 *   [1] stack_lds  LDS ring entries per lane (4, 5, 6, 8 or 16); deeper entries
 *                  spill to a per-lane area in HBM
 *   [2] leaf_ilp   triangle records loaded together per leaf step (1 or 2)
 *   [3] fetch_min  a persistent wave refills its finished lanes (from its
 *                  pool of ray indices) when at least this many are idle
 *                  (1..64)
 *   [4] occupancy  percent (1..100) of the waves that fit on the device that
 *                  the persistent grid launches: the workgroup count
 *   [5] wg_waves   waves per workgroup (1, 2 or 4): the workgroup size; the
 *                  waves of a workgroup trace independently
 *   [6] leaf_batch 0: every iteration steps the wave's node lanes and its
 *                  leaf lanes; T (1..64): an iteration steps only the leaf
 *                  lanes once at least T lanes wait at a leaf (or no lane
 *                  has node work), only the node lanes otherwise, so the
 *                  wave issues one of the two code paths per iteration
 *                  (packed layout only)
 *   [7] node_levels 1: a node step takes one tree level (one 16-byte
 *                  children pair); 2: it takes two levels from one 32-byte
 *                  record of the node and its children (one cache line per
 *                  two levels; packed layout, and only for trees whose
 *                  children pairs have equal clips -- the SAH trees -- whose
 *                  records the upload or GPU build makes; PRAY_ERR_UNSUPPORTED
 *                  for other trees); 3: it takes one level from a 64-byte
 *                  record of both children's boxes, over the boxed layout of
 *                  pray_boxed_layout (packed layout; every tree the upload
 *                  takes, PRAY_ERR_UNSUPPORTED for trees built on the
 *                  device); 4: it takes two binary levels from a 128-byte
 *                  record of up to four slots' boxes, over the layout's
 *                  4-wide form (the same trees; the upload's default)
 * pray_set_tuning: -1 keeps a value.  Calls of one context must not overlap
 * (traces share its ray counter and stack spill area). */
#define PRAY_TUNING_FIELDS 8
PRAY_API int pray_set_tuning(pray_ctx *ctx, int scene_id, const int *tuning);
PRAY_API int pray_get_tuning(pray_ctx *ctx, int scene_id, int *tuning);

/* Autotuner (no reference counterpart): times closest-hit traversal of
 * n_rays incoherent rays through the scene (0 = 2^20) over stack depth x
 * leaf ILP x refill threshold, then workgroup size and occupancy for the
 * fastest of those, and
 * keeps the fastest setting for the scene and for its size bucket (triangle
 * counts within a factor of 4).  Later scenes of a
 * tuned bucket take the bucket's setting without timing unless `force` is
 * set.  tuning (optional) receives the setting, PRAY_TUNING_FIELDS ints. */
PRAY_API int pray_autotune(pray_ctx *ctx, int scene_id, uint32_t n_rays, int force, int *tuning);

/* Autotuner for path tracing (no reference counterpart): the same search as
 * pray_autotune, timed on two-sample pray_render_pt renders of the caller's
 * image and settings, whose bounce queues trace more coherent rays than the
 * incoherent stream pray_autotune times.  Keeps the fastest setting for this
 * scene only (not for its size bucket).  tuning (optional) receives it. */
PRAY_API int pray_autotune_pt(pray_ctx *ctx, int scene_id, const float *camera16, uint32_t width, uint32_t height,
                              uint32_t max_depth, uint32_t rr_depth, uint64_t seed, const float *background,
                              float det_eps, int *tuning);

/* Closest-hit or any-hit query for n rays (replaces SpatialDatabase::intersect,
 * spatial.h:15, and the shadow tests of pray.cpp:39-44 / :123-133).
 *   rays     n x 6 floats: origin xyz, direction xyz (normalised)
 *   limit2   n floats (squared distance limit), ANY modes only, else NULL
 *   det_eps  determinant cull: a triangle is skipped when det < det_eps
 *            (the reference uses 1e-4, geometry.h:12)
 *   hit_idx  n int32: triangle index, -1 on a miss
 *   hit_t    n floats: distance, FLT_MAX on a miss (may be NULL) */
PRAY_API int pray_trace(pray_ctx *ctx, int scene_id, const float *rays, uint64_t n, int mode, float det_eps,
               const float *limit2, int32_t *hit_idx, float *hit_t);
PRAY_API int pray_trace_device(pray_ctx *ctx, int scene_id, const float *d_rays, uint64_t n, int mode, float det_eps,
                      const float *d_limit2, int32_t *d_hit_idx, float *d_hit_t, void *stream);

/* Same traversal, additionally writing per-ray work counts: d_counts[2*i] =
 * nodes visited (8-byte nodes; every child-pair fetch counts 2, the root 1),
 * d_counts[2*i+1] = leaf triangle references tested. */
PRAY_API int pray_trace_counts_device(pray_ctx *ctx, int scene_id, const float *d_rays, uint64_t n, int mode,
                             float det_eps, const float *d_limit2, int32_t *d_hit_idx, float *d_hit_t,
                             uint32_t *d_counts, void *stream);

/* pray_trace_device that allocates nothing.  It has no counterpart in the
 * reference; it is for callers that cannot allocate device memory while they
 * trace: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Same arguments and results as pray_trace_device (the
 * caller still passes rays and outputs; the scene is uploaded beforehand).
 * The chunk counters and the stack spill area live in a static array of
 * 64 MiB per GPU (device memory held from the library's load until the
 * process ends) instead of the context's buffers: when the spill slots of
 * the whole persistent grid do not fit, fewer waves trace the rays, with
 * identical results.  Rays per call: at most 2^31-1 (PRAY_ERR_INVALID
 * beyond); a tree whose spill does not fit one workgroup gives
 * PRAY_ERR_UNSUPPORTED (not reachable at max_depth <= 64).  The array keeps
 * the last call's contents and the counters are cleared by a kernel at the
 * start of every call; all calls on one GPU share it, from every context, so
 * calls that overlap on one GPU (other streams, other contexts or threads)
 * race and must be serialised by the caller. */
PRAY_API int pray_trace_device_noalloc(pray_ctx *ctx, int scene_id, const float *d_rays, uint64_t n, int mode,
                                       float det_eps, const float *d_limit2, int32_t *d_hit_idx, float *d_hit_t,
                                       void *stream);

/* Incoherent ray stream: ray i has origin lo + span * U and a direction
 * uniform on the sphere (z = 1 - 2u, phi = 2 pi u'), normalised as the
 * reference normalises; U and u come from Philox4x32-10 keyed by `seed`
 * with counters (i, 0, 0, 0) and (i, 1, 0, 0).  box6 = lo xyz, span xyz. */
PRAY_API int pray_gen_rays_device(pray_ctx *ctx, uint64_t seed, const float *box6, uint64_t n, float *d_rays, void *stream);
PRAY_API int pray_gen_rays(pray_ctx *ctx, uint64_t seed, const float *box6, uint64_t n, float *rays);

/* Camera of the reference (scene_preprocessor.cpp:98-136) as 16 floats:
 * position xyz, inverse view matrix rows 0..2 x cols 0..2 (9), angle,
 * 1/width, 1/height, width/height.  Writes width*height primary rays
 * (n x 6 floats, row-major pixels, top row first). */
PRAY_API int pray_camera_rays(pray_ctx *ctx, const float *camera16, uint32_t width, uint32_t height, float *rays);

/* Whitted ray tracing of one image (replaces rayTrace, pray.cpp:380-590):
 * primary rays from the camera, closest hit, per-light Lambert with an any-hit
 * shadow ray (inclusive limit, pray.cpp:129), clamp and round to RGB8.
 *   lights    n_lights x 6 floats: position xyz, colour xyz
 *   background 3 floats
 *   rgb       width*height*3 bytes (top row first)
 *   hit_idx   width*height int32 primary hits, may be NULL */
PRAY_API int pray_render_rt(pray_ctx *ctx, int scene_id, const float *camera16, uint32_t width, uint32_t height,
                   const float *lights, uint32_t n_lights, const float *background, float det_eps, uint8_t *rgb,
                   int32_t *hit_idx);

/* Wavefront path tracing (replaces pathTraceOnGPURecursive,
 * path_trace.cu:494-555, NORMAL mode estimator pray.cpp:342-378):
 * per sample pass, generate -> (extend -> shade)* over HBM ray queues,
 * Philox4x32-10 keyed by (seed, pixel, sample, bounce); Russian roulette
 * from bounce `rr_depth` on (rr_depth >= max_depth disables it).  At
 * max_depth <= 1 the reference's DIFFUSE estimator instead (pray.cpp:315-340,
 * chosen there by scene_preprocessor.cpp:34-40, also for pray_render_pt_batch):
 * a camera-ray miss gives the background and an emissive hit its emission,
 * exactly; otherwise the samples' arriving radiance is summed and multiplied
 * by the hit's albedo once, (sum * albedo) / samples; no roulette.
 *   hdr   width*height*3 floats: mean radiance before clamping, may be NULL
 *   rgb   width*height*3 bytes: clamp(mean)*255+0.5 truncated, may be NULL */
PRAY_API int pray_render_pt(pray_ctx *ctx, int scene_id, const float *camera16, uint32_t width, uint32_t height,
                   uint32_t spp, uint32_t max_depth, uint32_t rr_depth, uint64_t seed, const float *background,
                   float det_eps, float *hdr, uint8_t *rgb);

/* pray_render_pt with the sample pass either captured once into a hipGraph
 * and replayed spp times (use_graph = 1, the default of pray_render_pt) or
 * enqueued kernel by kernel (use_graph = 0); results are identical. */
PRAY_API int pray_render_pt_ex(pray_ctx *ctx, int scene_id, const float *camera16, uint32_t width, uint32_t height,
                               uint32_t spp, uint32_t max_depth, uint32_t rr_depth, uint64_t seed,
                               const float *background, float det_eps, float *hdr, uint8_t *rgb, int use_graph);

/* Progressive path tracing (the reference renders all samples in one call,
 * path_trace.cu:494-555, and keeps only the clamped 8-bit image,
 * pray.cpp:377 / main.cpp:93-98): samples first_sample .. first_sample+spp-1
 * of the image pray_render_pt would render with the same arguments.
 *   accum  width*height*3 floats, in/out: the running sum of per-sample
 *          radiance over samples 0 .. first_sample-1 (not read when
 *          first_sample = 0); on return the sum over 0 .. first_sample+spp-1
 *   hdr / rgb  as pray_render_pt, the mean over all first_sample+spp samples
 * Renders split over several calls (passing accum on) give the image of one
 * pray_render_pt call bit for bit.  At most 2^24 samples in total.
 * The path tracers keep their image-sized device buffers in the context
 * (about 170 bytes per pixel for pray_render_pt*, 240 for the cache tracer)
 * and reuse them in later calls until pray_destroy. */
PRAY_API int pray_render_pt_progressive(pray_ctx *ctx, int scene_id, const float *camera16, uint32_t width,
                                        uint32_t height, uint32_t first_sample, uint32_t spp, uint32_t max_depth,
                                        uint32_t rr_depth, uint64_t seed, const float *background, float det_eps,
                                        float *accum, float *hdr, uint8_t *rgb);

/* Path tracing of several scenes in one wavefront (no reference
 * counterpart): every bounce is one persistent closest-hit launch over the
 * paths of all scenes (rays carry their scene; packed layout) and one
 * shading launch; the sample pass is captured once into a hipGraph and
 * replayed spp times.  Scene k renders widths[k] x heights[k] pixels with
 * camera cameras16[16k..], seed seeds[k] and background backgrounds[3k..];
 * its image equals pray_render_pt's for the same arguments bit for bit.  The
 * launches use the setting of the scene with the largest image in the size
 * bucket holding most pixels: its bucket's after pray_autotune, its own after
 * pray_autotune_pt or pray_set_tuning, the defaults otherwise.  hdr[k] / rgb[k] (arrays of per-scene output
 * pointers, each may be NULL); gpu_ms (optional) receives the device time of
 * the spp passes. */
PRAY_API int pray_render_pt_batch(pray_ctx *ctx, uint32_t n_scenes, const int *scene_ids, const float *cameras16,
                                  const uint32_t *widths, const uint32_t *heights, uint32_t spp, uint32_t max_depth,
                                  uint32_t rr_depth, const uint64_t *seeds, const float *backgrounds, float det_eps,
                                  float *const *hdr, uint8_t *const *rgb, float *gpu_ms);
/* Rays traced (all bounces, all samples) by the last pray_render_pt* or
 * pray_render_pt_batch call of the context.  Every sample of a pixel starts
 * with the pixel's camera ray, which a call traces once and whose hit its
 * samples share: the camera rays count once per call. */
PRAY_API int pray_last_render_rays(pray_ctx *ctx, uint64_t *rays);
/* Per bounce b (0 = camera rays), the paths at that bounce summed over all
 * samples of the last pray_render_pt / _ex / _progressive call:
 * the sizes of the wavefront queues.  Fills min(n, *n_bounces) counts;
 * *n_bounces = max_depth + 1 (0 after other calls). */
PRAY_API int pray_last_render_queues(pray_ctx *ctx, uint64_t *counts, uint32_t n, uint32_t *n_bounces);

/* Learned radiance cache (no reference counterpart; the reference's path
 * tracer, path_trace.cu:369-464, always traces paths to the end).
 *
 * An MLP predicts the radiance arriving at a surface point, divided by the
 * point's albedo, from 10 features: position in the scene box mapped to
 * [0,1]^3, ray direction and surface normal as (theta/pi, phi/2pi + 1/2),
 * albedo rgb.  Network: 12-frequency sin/cos encoding of the 7 geometric
 * features + albedo + a constant 1 (172 of 192 inputs) -> 5 hidden layers of
 * width 64 with ReLU -> 3 outputs; BF16 weights and activations with FP32
 * accumulation on the matrix cores, FP32 master weights and Adam (beta1 0.9,
 * beta2 0.99, eps 1e-8).  Loss: relative L2, mean over samples and channels
 * of (y - t)^2 / (y^2 + 0.01) with the denominator held constant.
 * Parameters are PRAY_CACHE_PARAMS floats, row-major [out][in] per layer:
 * 64x192, 4 x 64x64, 32x64; rows 3.. of the last layer and the 20 input
 * channels of the first layer that the encoding leaves at zero are padding
 * and stay zero.  The channel order of the encoding is defined in
 * nostrum_pray_amd/csrc/mlp.hpp (enc_value). */
#define PRAY_CACHE_PARAMS 30720
#define PRAY_CACHE_FEATURES 10
PRAY_API int pray_cache_create(pray_ctx *ctx, uint64_t seed, float learning_rate, uint32_t max_batch, int *cache_id);
PRAY_API int pray_cache_destroy(pray_ctx *ctx, int cache_id);
PRAY_API int pray_cache_get_params(pray_ctx *ctx, int cache_id, float *params);
/* Sets the master weights and resets the optimiser state. */
PRAY_API int pray_cache_set_params(pray_ctx *ctx, int cache_id, const float *params);
/* out: n x 3 predictions for n x 10 features. */
PRAY_API int pray_cache_infer(pray_ctx *ctx, int cache_id, const float *features, uint64_t n, float *out);
/* Gradient of the batch loss (n <= max_batch samples, targets n x 3) without
 * an update; grad: PRAY_CACHE_PARAMS floats; loss optional. */
PRAY_API int pray_cache_gradient(pray_ctx *ctx, int cache_id, const float *features, const float *targets, uint64_t n,
                                 float *grad, float *loss);
/* One Adam step on a batch; loss (before the step) optional. */
PRAY_API int pray_cache_train_step(pray_ctx *ctx, int cache_id, const float *features, const float *targets,
                                   uint64_t n, float *loss);
PRAY_API int pray_cache_stats(pray_ctx *ctx, int cache_id, uint32_t *steps, float *loss_last, float *loss_running,
                              int *cache_on);

/* Path tracing with the radiance cache, `frames` samples per pixel.  Each
 * frame (one hipGraph replay) traces one path per pixel as pray_render_pt
 * does, except at a non-emissive hit of bounce `cache_bounce` (0 = the
 * primary hit):
 *   - while the cache is on, the path stops there and its pixel receives
 *     throughput * albedo * max(cache prediction, 0);
 *   - at `train_paths` training pixels per frame (every stride-th pixel from
 *     an offset drawn per frame) the hit's features become a training sample
 *     and one more ray leaves it (own random stream); the target is that
 *     ray's emission or background, or albedo * the cache's prediction at its
 *     hit.  One Adam step on the frame's samples follows.
 * The cache is on once at least `min_steps` steps were taken and the running
 * training loss is below `max_loss`; otherwise the frame is plain path tracing
 * with pray_render_pt's random numbers, so max_loss = 0 gives exactly
 * pray_render_pt's image.  frame_stats (optional): frames x {training loss,
 * cache on}. */
PRAY_API int pray_render_pt_cache(pray_ctx *ctx, int scene_id, int cache_id, const float *camera16, uint32_t width,
                                  uint32_t height, uint32_t frames, uint32_t max_depth, uint32_t rr_depth,
                                  uint64_t seed, const float *background, float det_eps, uint32_t train_paths,
                                  uint32_t cache_bounce, float max_loss, uint32_t min_steps, float *hdr, uint8_t *rgb,
                                  float *frame_stats);
/* As pray_render_pt_cache with `samples_per_frame` paths per pixel in each
 * frame (frames x samples_per_frame samples in all, traced together): sample
 * s of a frame is sample frame * samples_per_frame + s of pray_render_pt's
 * random streams; training pixels come from each frame's first sample, so a
 * frame still takes one Adam step on at most `train_paths` samples.  With
 * max_loss = 0 the image is pray_render_pt's at frames x samples_per_frame
 * samples per pixel.  samples_per_frame = 1 is pray_render_pt_cache. */
PRAY_API int pray_render_pt_cache_ex(pray_ctx *ctx, int scene_id, int cache_id, const float *camera16,
                                     uint32_t width, uint32_t height, uint32_t frames, uint32_t samples_per_frame,
                                     uint32_t max_depth, uint32_t rr_depth, uint64_t seed, const float *background,
                                     float det_eps, uint32_t train_paths, uint32_t cache_bounce, float max_loss,
                                     uint32_t min_steps, float *hdr, uint8_t *rgb, float *frame_stats);

/* Device-memory and timing utilities (no reference counterpart: the
 * reference allocates and copies inside every call).  They let a caller keep
 * rays and hit records resident in HBM between calls and time work on the
 * context's stream with HIP events. */
PRAY_API int pray_device_alloc(pray_ctx *ctx, uint64_t bytes, void **d_ptr);
PRAY_API int pray_device_free(pray_ctx *ctx, void *d_ptr);
/* Page-locked host memory (no reference counterpart) for results that every
 * call writes back, such as a frame loop's image and hit buffers: the device
 * copies into it directly, while a copy into pageable memory passes through
 * the runtime's staging buffers.  pray_host_free takes no context, so memory
 * may outlive the context that allocated it. */
PRAY_API int pray_host_alloc(pray_ctx *ctx, uint64_t bytes, void **ptr);
PRAY_API int pray_host_free(void *ptr);
PRAY_API int pray_copy_to_device(pray_ctx *ctx, void *d_dst, const void *src, uint64_t bytes);
PRAY_API int pray_copy_to_host(pray_ctx *ctx, void *dst, const void *d_src, uint64_t bytes);
PRAY_API int pray_sync(pray_ctx *ctx);
/* Device-to-device copy by a plain streaming kernel (16-byte loads/stores);
 * bytes must be a multiple of 16.  Also the reference point for calibrating
 * HBM traffic counters. */
PRAY_API int pray_copy_device(pray_ctx *ctx, void *d_dst, const void *d_src, uint64_t bytes, void *stream);
/* Records an event on the context's stream; slot 0..15. */
PRAY_API int pray_event_record(pray_ctx *ctx, int slot);
/* Milliseconds between two recorded events (synchronises on the second). */
PRAY_API int pray_event_elapsed(pray_ctx *ctx, int slot0, int slot1, float *ms);
/* With enable = 0, the path tracers (pray_render_pt*, pray_render_pt_cache,
 * pray_render_pt_batch) enqueue every sample pass kernel by kernel instead of
 * replaying a hipGraph captured once per call; results are identical.  For
 * tools that follow kernel launches one by one (profilers); the default is 1. */
PRAY_API int pray_use_graphs(pray_ctx *ctx, int enable);
/* Paths the plain path tracers (pray_render_pt, pray_render_pt_ex,
 * pray_render_pt_progressive) keep in flight: a call traces
 * S = min(spp, max(1, max_paths / pixels)) samples of every pixel together,
 * so each bounce's trace launch holds up to S x pixels paths, in passes of S
 * samples (the last pass takes the rest).  Images, running sums and ray
 * counts do not depend on S.  Device memory of a call grows by about
 * 100 bytes per path in flight; the default is 2^24 paths (1.6 GB), and 0
 * traces one sample per pass.  S is also capped so that the path buffers
 * stay within half the device's free memory at the call.  No counterpart in
 * the reference, whose kernel traces one sample per pixel launch. */
PRAY_API int pray_set_path_budget(pray_ctx *ctx, uint64_t max_paths);
/* Kernel timing: with enable = 1, every later trace that is not being
 * captured into a graph records events around its traversal kernel alone
 * (not the counter reset before it), on the context's stream, without
 * synchronising; the events of the last 64 such kernels are kept.
 * Switching timing on (from off) forgets earlier measurements.
 * pray_last_kernel_ms synchronises and returns the last such kernel's
 * duration; pray_timed_kernels_ms synchronises and returns the mean
 * duration of the last min(64, timed) kernels and their number.  Traces
 * enqueued back to back and read once at the end give the kernel's
 * duration with its launches pipelined (no host round trip per launch). */
PRAY_API int pray_kernel_timing(pray_ctx *ctx, int enable);
PRAY_API int pray_last_kernel_ms(pray_ctx *ctx, float *ms);
PRAY_API int pray_timed_kernels_ms(pray_ctx *ctx, float *mean_ms, uint32_t *count);

#ifdef __cplusplus
}
#endif

#endif /* PRAY_HIP_H */
