/*
 * pray_host.h -- C ABI of libpray_host.so, the CPU side of the Nostrum Pray
 * drop-in: scene loading and procedural scenes, the median and SAH KD-tree
 * builders (which stay on the CPU) and the flattening into the 8-byte-node
 * device layout.  Callers: the Pray CLI, the Python tests and bench.py.
 *
 *   ph_scene_load_json   reference readSceneDescription (include/io.h:9)
 *   ph_build             reference preprocessScene tree build
 *                        (src/scene_preprocessor.cpp:52-70), KDTree::build
 *                        (src/kd_tree.cpp:25-46), KDTreeSAH::build
 *                        (src/kd_tree_sah.cpp:32-90)
 *   ph_flatten           reference buildGPUTreeFromKDTree (include/gpu_util.h:10)
 *   ph_camera / ph_prim_rays   reference calculatePrimRays
 *                        (src/scene_preprocessor.cpp:98-136)
 *   ph_upload            flatten + pray_upload_scene in one call
 *
 * Handles are opaque pointers; all functions returning int use 0 for success.
 */
#ifndef PRAY_HOST_H
#define PRAY_HOST_H

#include <stdint.h>

#ifndef PRAY_API
#define PRAY_API __attribute__((visibility("default")))
#endif

#include "pray_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ph_scene ph_scene;
typedef struct ph_tree ph_tree;

/* Procedural scenes (see nostrum_pray_amd/host/scene.hpp). */
PRAY_API ph_scene *ph_scene_terrain(uint32_t grid, uint64_t seed);
PRAY_API ph_scene *ph_scene_spheres(uint64_t seed);
PRAY_API ph_scene *ph_scene_cornell(uint64_t seed, uint32_t target_tris);
PRAY_API ph_scene *ph_scene_random(uint64_t seed, uint32_t n_tris);
/* Builds a scene from raw triangles: abc = n x 9 floats (a, b, c), color =
 * n x 3 floats, emit = n bytes (may be NULL). */
PRAY_API ph_scene *ph_scene_from_tris(const float *abc, const float *color, const uint8_t *emit, uint64_t n);
/* Reference scene JSON + OBJ; NULL on failure with the message in err. */
PRAY_API ph_scene *ph_scene_load_json(const char *path, char *err, uint64_t err_len);
PRAY_API int ph_scene_write_json(const ph_scene *s, const char *json_path);
PRAY_API void ph_scene_free(ph_scene *s);

PRAY_API uint64_t ph_scene_num_tris(const ph_scene *s);
/* abc: n x 9 raw vertices, color n x 3, emit n (any may be NULL). */
PRAY_API void ph_scene_get_tris(const ph_scene *s, float *abc, float *color, uint8_t *emit);
/* Precomputed fields after initialiseTriangles: ab, ac, normal (n x 3 each;
 * any may be NULL). */
PRAY_API void ph_scene_get_derived(const ph_scene *s, float *ab, float *ac, float *normal);
/* info[16]: mode, width, height, fov, max_depth, num_samples, n_lights,
 * bg r g b, det_eps, rr_depth (-1 = off), seed, tree kind (-1 auto), 0, 0 */
PRAY_API void ph_scene_info(const ph_scene *s, double *info);
/* Overrides width/height/spp/depth (0 keeps a value; rr_depth < 0 keeps). */
PRAY_API void ph_scene_set_render(ph_scene *s, uint32_t width, uint32_t height, uint32_t spp, uint32_t max_depth, int rr_depth);
/* lights: n_lights x 6 floats (position, colour). */
PRAY_API void ph_scene_get_lights(const ph_scene *s, float *lights);
/* camera16 as taken by pray_render_* / pray_camera_rays. */
PRAY_API void ph_camera(const ph_scene *s, float *camera16);
/* width*height*6 floats; host restatement of calculatePrimRays. */
PRAY_API void ph_prim_rays(const ph_scene *s, float *rays);

/* Tree build over the scene's triangles: kind 0 = median, 1 = SAH (the
 * reference's event sweep), 2 = binned SAH (64 bins per axis, this build's
 * extension: the CPU twin of pray_build_scene's PRAY_BUILD_SAH, flattened
 * with the host's treelet layout), -1 = the reference's policy. */
PRAY_API ph_tree *ph_build(ph_scene *s, int kind);
/* Same with an explicit depth cap (the reference's is 20; deeper trees have
 * smaller leaves and produce identical hits). */
PRAY_API ph_tree *ph_build_depth(ph_scene *s, int kind, int max_depth);
/* Same with the leaf size below which nodes are not split (reference 32,
 * kd_tree.h:50) and, for the SAH trees, the traversal and intersection costs
 * (reference 5 and 1.5, kd_tree_sah.h:49-50; binned SAH 2 and 1).  Values
 * <= 0 keep the defaults.  Any tree produces the same closest hits; these only trade
 * node visits against triangle tests. */
PRAY_API ph_tree *ph_build_ex(ph_scene *s, int kind, int max_depth, int leaf_size, float traversal_cost,
                              float intersection_cost);
PRAY_API void ph_tree_free(ph_tree *t);
/* Flat-tree file (no reference counterpart): the device layout, counts and
 * root box, checksummed, so that a large scene is built once and later
 * uploaded without a rebuild.  A loaded tree has no CPU tree (ph_tree_export
 * reports 0 nodes); its stats come from the flat layout.  Return 0 / NULL on
 * failure with the reason in err. */
PRAY_API int ph_tree_save(const ph_tree *t, const char *path, char *err, uint64_t err_len);
PRAY_API ph_tree *ph_tree_load(const char *path, char *err, uint64_t err_len);
/* stats[8]: kind, nnodes, maxdepth, leafNodes, emptyLeafNodes, triangleCount,
 * flat node count, flat leaf refs */
PRAY_API void ph_tree_stats(const ph_tree *t, uint64_t *stats);
/* Depth-first export (left before right).  Call with NULL buffers to get
 * sizes[2] = {nodes, leaf refs}; then node_i32 = nodes x 4
 * {is_leaf, axis, first, count}, node_f32 = nodes x 7 {split, box[6]},
 * tris = leaf refs. */
PRAY_API void ph_tree_export(const ph_tree *t, uint64_t *sizes, int32_t *node_i32, float *node_f32, uint32_t *tris);
/* Flat layout: sizes[4] = {n_nodes, n_refs, n_tris, max_depth}; buffers may
 * be NULL. */
PRAY_API void ph_flat_get(const ph_tree *t, uint64_t *sizes, uint32_t *nodes, uint32_t *refs, float *a, float *ab, float *ac,
                 float *normal, float *color, uint8_t *emit, float *root_box);
/* Float images and progressive-render state (no reference counterpart: the
 * reference writes only the clamped 8-bit BMP, main.cpp:93-98).
 * ph_write_pfm / ph_read_pfm: RGB Portable Float Map, w*h*3 floats top row
 * first (ph_read_pfm with rgb = NULL only reports wh[2] = {w, h}).
 * ph_scene_fingerprint: hash of everything that decides a path-traced pixel
 * except the sample count (triangles, materials, camera, size, depth limits,
 * seed, background, det_eps).
 * ph_accum_save / ph_accum_load: the running radiance sum of
 * pray_render_pt_progressive (w*h*3 floats) after `samples` samples, with the
 * fingerprint and a checksum; info3 = {w, h, samples}; ph_accum_load
 * checks the whole file and with sum = NULL only reports info3 and the
 * fingerprint.  Return 0, or -1 with the reason in err. */
PRAY_API uint64_t ph_scene_fingerprint(const ph_scene *s);
PRAY_API int ph_write_pfm(const char *path, uint32_t w, uint32_t h, const float *rgb);
PRAY_API int ph_read_pfm(const char *path, uint32_t *wh, float *rgb, char *err, uint64_t err_len);
PRAY_API int ph_accum_save(const char *path, uint32_t w, uint32_t h, uint32_t samples, uint64_t fingerprint,
                           const float *sum, char *err, uint64_t err_len);
PRAY_API int ph_accum_load(const char *path, uint32_t *info3, uint64_t *fingerprint, float *sum, char *err,
                           uint64_t err_len);
/* Uploads the flat layout into a pray context; with_shading = 0 skips
 * normals/colours/emission. */
PRAY_API int ph_upload(const ph_tree *t, pray_ctx *ctx, int with_shading, int *scene_id);

#ifdef __cplusplus
}
#endif

#endif /* PRAY_HOST_H */
