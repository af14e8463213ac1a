// ORACLE -- test infrastructure only.
//
// C entry points over the reference's own classes (compiled from the
// sources under /root/reference by oracle/Makefile into oracle/_ref/).  The
// reference API returns Triangle copies, so each triangle carries its index
// in its colour field (x = idx & 0xFFF, y = idx >> 12), which the
// intersection code never reads.  Rays are built without the normalising Ray
// constructor: callers pass normalised directions, exactly as the reference's
// own rays hold them after construction.
#include <omp.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "bf_search.h"
#include "io.h"
#include "kd_tree.h"
#include "kd_tree_sah.h"
#include "pray.h"
#include "scene_preprocessor.h"

#define REF_API extern "C" __attribute__((visibility("default")))

namespace {
struct RefDB {
    std::unique_ptr<pray::SpatialDatabase> db;
};

inline pray::Ray mkRay(const float *r) {
    pray::Ray ray;
    ray.position = pray::Vec3(r[0], r[1], r[2]);
    ray.direction = pray::Vec3(r[3], r[4], r[5]);
    ray.inverseDirection = ray.direction.inverse();
    return ray;
}

inline uint32_t idxOf(const pray::Triangle &t) {
    return (uint32_t)t.color.x | ((uint32_t)t.color.y << 12);
}

void exportRec(const pray::KDTreeElement *n, std::vector<int32_t> &ni, std::vector<float> &nf,
               std::vector<uint32_t> &tris) {
    const bool leaf = !n->left && !n->right;
    ni.push_back(leaf ? 1 : 0);
    ni.push_back(n->splitPlane.axis);
    ni.push_back((int32_t)tris.size());
    ni.push_back((int32_t)n->triangleIndices.size());
    nf.push_back(n->splitPlane.distance);
    const float b[6] = {n->boundingBox.min.x, n->boundingBox.min.y, n->boundingBox.min.z,
                        n->boundingBox.max.x, n->boundingBox.max.y, n->boundingBox.max.z};
    nf.insert(nf.end(), b, b + 6);
    for (size_t t : n->triangleIndices) tris.push_back((uint32_t)t);
    if (!leaf) {
        if (n->left) exportRec(n->left.get(), ni, nf, tris);
        if (n->right) exportRec(n->right.get(), ni, nf, tris);
    }
}
}  // namespace

// kind 0 = KDTree (median), 1 = KDTreeSAH, 2 = BFSearch.
REF_API void *ref_build(const float *abc, uint64_t n, int kind) {
    pray::Scene scene;
    scene.triangles.reserve(n);
    for (uint64_t i = 0; i < n; i++) {
        const float *v = abc + 9 * i;
        const pray::Vec3 col((float)(i & 0xFFF), (float)(i >> 12), 0.f);
        scene.triangles.emplace_back(pray::Vec3(v[0], v[1], v[2]), pray::Vec3(v[3], v[4], v[5]),
                                     pray::Vec3(v[6], v[7], v[8]), col);
    }
    pray::initialiseTriangles(scene);
    auto *r = new RefDB();
    if (kind == 0)
        r->db = std::make_unique<pray::KDTree>();
    else if (kind == 1)
        r->db = std::make_unique<pray::KDTreeSAH>();
    else
        r->db = std::make_unique<pray::BFSearch>();
    r->db->build(std::move(scene.triangles));
    return r;
}

REF_API void ref_free(void *h) { delete static_cast<RefDB *>(h); }

REF_API void ref_stats(void *h, uint64_t *o) {
    auto *r = static_cast<RefDB *>(h);
    o[0] = r->db->nnodes, o[1] = r->db->maxdepth, o[2] = r->db->leafNodes, o[3] = r->db->emptyLeafNodes,
    o[4] = r->db->triangleCount;
}

// Scalar closest hit through SpatialDatabase::intersect.
REF_API void ref_intersect(void *h, const float *rays, uint64_t n, int32_t *idx, float *t) {
    auto *r = static_cast<RefDB *>(h);
#pragma omp parallel for schedule(dynamic, 16)
    for (uint64_t i = 0; i < n; i++) {
        pray::Triangle tri;
        float d = 0;
        const bool hit = r->db->intersect(mkRay(rays + 6 * i), tri, d);
        idx[i] = hit ? (int32_t)idxOf(tri) : -1;
        t[i] = hit ? d : 0.f;
    }
}

// AVX packet closest hit (8 rays per call), OpenMP dynamic,10 over packets as
// rayTrace schedules them (pray.cpp:386).  Returns the traversal wall time.
REF_API double ref_intersect_avx(void *h, const float *rays, uint64_t n, int threads, int32_t *idx, float *t) {
    auto *r = static_cast<RefDB *>(h);
    const uint64_t packets = (n + 7) / 8;
    const double t0 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 10) num_threads(threads)
    for (uint64_t p = 0; p < packets; p++) {
        alignas(32) float c[6][8];
        for (int j = 0; j < 8; j++) {
            const uint64_t k = 8 * p + j < n ? 8 * p + j : n - 1;
            for (int q = 0; q < 6; q++) c[q][j] = rays[6 * k + q];
        }
        std::array<pray::Triangle, 8> trs;
        __m256 dist;
        const __m256 m = r->db->intersect(_mm256_load_ps(c[0]), _mm256_load_ps(c[1]), _mm256_load_ps(c[2]),
                                          _mm256_load_ps(c[3]), _mm256_load_ps(c[4]), _mm256_load_ps(c[5]),
                                          _mm256_castsi256_ps(_mm256_set1_epi32(-1)), trs, dist);
        alignas(32) float dv[8];
        _mm256_store_ps(dv, dist);
        const int mm = _mm256_movemask_ps(m);
        for (int j = 0; j < 8 && 8 * p + j < n; j++) {
            const bool hit = (mm >> j) & 1;
            idx[8 * p + j] = hit ? (int32_t)idxOf(trs[j]) : -1;
            t[8 * p + j] = hit ? dv[j] : 0.f;
        }
    }
    return omp_get_wtime() - t0;
}

// Preorder export of a KD tree (same format as ph_tree_export); sizes[2].
REF_API void ref_export(void *h, uint64_t *sizes, int32_t *ni, float *nf, uint32_t *tris) {
    auto *r = static_cast<RefDB *>(h);
    auto *kd = dynamic_cast<pray::KDTree *>(r->db.get());
    std::vector<int32_t> vi;
    std::vector<float> vf;
    std::vector<uint32_t> vt;
    if (kd && kd->root) exportRec(kd->root.get(), vi, vf, vt);
    if (sizes) sizes[0] = vi.size() / 4, sizes[1] = vt.size();
    if (ni) std::memcpy(ni, vi.data(), vi.size() * 4);
    if (nf) std::memcpy(nf, vf.data(), vf.size() * 4);
    if (tris) std::memcpy(tris, vt.data(), vt.size() * 4);
}

// initialiseTriangles: ab, ac, normal (n x 3 each).
REF_API void ref_init_tris(const float *abc, uint64_t n, float *ab, float *ac, float *normal) {
    pray::Scene scene;
    for (uint64_t i = 0; i < n; i++) {
        const float *v = abc + 9 * i;
        scene.triangles.emplace_back(pray::Vec3(v[0], v[1], v[2]), pray::Vec3(v[3], v[4], v[5]),
                                     pray::Vec3(v[6], v[7], v[8]), pray::Vec3(0, 0, 0));
    }
    pray::initialiseTriangles(scene);
    for (uint64_t i = 0; i < n; i++) {
        const pray::Triangle &t = scene.triangles[i];
        const float e[9] = {t.ab.x, t.ab.y, t.ab.z, t.ac.x, t.ac.y, t.ac.z, t.normal.x, t.normal.y, t.normal.z};
        std::memcpy(ab + 3 * i, e, 12);
        std::memcpy(ac + 3 * i, e + 3, 12);
        std::memcpy(normal + 3 * i, e + 6, 12);
    }
}

// readSceneDescription (io.cpp:27-137): returns a handle holding the scene.
// info[12] = mode, width, height, fov, max_depth, num_samples, n_lights,
// n_tris, camera pos (3) ... see ref_scene_get.
REF_API void *ref_load_scene(const char *path) {
    auto *s = new pray::Scene();
    pray::readSceneDescription(path, *s);
    return s;
}

REF_API void ref_scene_free(void *h) { delete static_cast<pray::Scene *>(h); }

// info[16]: mode, width, height, fov, max_depth, num_samples, n_lights, n_tris,
// cam pos xyz, cam look xyz, 0, 0; tris n x 9, colors n x 3, emit n,
// lights n x 6, background 3 (any buffer may be NULL).
REF_API void ref_scene_get(void *h, double *info, float *tris, float *colors, uint8_t *emit, float *lights,
                           float *bg) {
    auto *s = static_cast<pray::Scene *>(h);
    if (info) {
        const double v[16] = {(double)s->mode, (double)s->width, (double)s->height, (double)s->field_of_view,
                              (double)s->maxDepth, (double)s->numSamples, (double)s->lights.size(),
                              (double)s->triangles.size(), s->camera_position.x, s->camera_position.y,
                              s->camera_position.z, s->camera_look.x, s->camera_look.y, s->camera_look.z, 0, 0};
        std::memcpy(info, v, sizeof v);
    }
    for (size_t i = 0; i < s->triangles.size(); i++) {
        const pray::Triangle &t = s->triangles[i];
        if (tris) {
            const float v[9] = {t.a.x, t.a.y, t.a.z, t.b.x, t.b.y, t.b.z, t.c.x, t.c.y, t.c.z};
            std::memcpy(tris + 9 * i, v, sizeof v);
        }
        if (colors) colors[3 * i] = t.color.x, colors[3 * i + 1] = t.color.y, colors[3 * i + 2] = t.color.z;
        if (emit) emit[i] = t.isEmitting ? 1 : 0;
    }
    if (lights)
        for (size_t i = 0; i < s->lights.size(); i++) {
            const pray::LightSource &l = s->lights[i];
            const float v[6] = {l.position.x, l.position.y, l.position.z, l.color.x, l.color.y, l.color.z};
            std::memcpy(lights + 6 * i, v, sizeof v);
        }
    if (bg) bg[0] = s->backgroundColor.x, bg[1] = s->backgroundColor.y, bg[2] = s->backgroundColor.z;
}

// calculatePrimRays: width*height rays, 6 floats each.
REF_API void ref_prim_rays(const float *pos, const float *look, uint64_t fov, uint64_t w, uint64_t h, float *rays) {
    pray::Scene scene;
    scene.camera_position = pray::Vec3(pos[0], pos[1], pos[2]);
    scene.camera_look = pray::Vec3(look[0], look[1], look[2]);
    scene.field_of_view = fov;
    scene.width = w;
    scene.height = h;
    pray::calculatePrimRays(scene);
    for (uint64_t i = 0; i < w * h; i++) {
        const pray::Ray &r = scene.primRays[i];
        const float v[6] = {r.position.x, r.position.y, r.position.z, r.direction.x, r.direction.y, r.direction.z};
        std::memcpy(rays + 6 * i, v, sizeof v);
    }
}
