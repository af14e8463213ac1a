// extern "C" surface of libpray_host.so (include/pray_host.h).
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pray_host.h"
#include "gpu_layout.hpp"
#include "scene.hpp"

struct ph_scene {
    pray::Scene s;
    bool derived = false;
    void ensureDerived() {
        if (!derived) {
            pray::initialiseTriangles(s);
            derived = true;
        }
    }
};

struct ph_tree {
    int kind = 0;
    std::unique_ptr<pray::KDTree> tree;
    pray::KDTreeGPU flat;
};

extern "C" {

ph_scene *ph_scene_terrain(uint32_t grid, uint64_t seed) {
    auto *p = new ph_scene();
    pray::makeTerrainScene(p->s, grid, seed);
    return p;
}

ph_scene *ph_scene_spheres(uint64_t seed) {
    auto *p = new ph_scene();
    pray::makeSpheresScene(p->s, seed);
    return p;
}

ph_scene *ph_scene_cornell(uint64_t seed, uint32_t target_tris) {
    auto *p = new ph_scene();
    pray::makeCornellScene(p->s, seed, target_tris);
    return p;
}

ph_scene *ph_scene_random(uint64_t seed, uint32_t n_tris) {
    auto *p = new ph_scene();
    pray::makeRandomScene(p->s, seed, n_tris);
    return p;
}

ph_scene *ph_scene_from_tris(const float *abc, const float *color, const uint8_t *emit, uint64_t n) {
    auto *p = new ph_scene();
    p->s.triangles.reserve(n);
    for (uint64_t i = 0; i < n; i++) {
        const float *v = abc + 9 * i;
        const pray::Vec3 col = color ? pray::Vec3(color[3 * i], color[3 * i + 1], color[3 * i + 2])
                                     : pray::Vec3(1.f, 1.f, 1.f);
        p->s.triangles.emplace_back(pray::Vec3(v[0], v[1], v[2]), pray::Vec3(v[3], v[4], v[5]),
                                    pray::Vec3(v[6], v[7], v[8]), col, emit ? emit[i] != 0 : false);
    }
    return p;
}

ph_scene *ph_scene_load_json(const char *path, char *err, uint64_t err_len) {
    auto *p = new ph_scene();
    std::string e;
    if (!pray::readSceneDescription(path, p->s, e)) {
        if (err && err_len) {
            std::strncpy(err, e.c_str(), err_len - 1);
            err[err_len - 1] = '\0';
        }
        delete p;
        return nullptr;
    }
    return p;
}

int ph_scene_write_json(const ph_scene *s, const char *json_path) {
    std::string e;
    return pray::writeSceneFiles(s->s, json_path, e) ? 0 : -1;
}

void ph_scene_free(ph_scene *s) { delete s; }

uint64_t ph_scene_num_tris(const ph_scene *s) { return s->s.triangles.size(); }

void ph_scene_get_tris(const ph_scene *s, float *abc, float *color, uint8_t *emit) {
    const auto &T = s->s.triangles;
    for (size_t i = 0; i < T.size(); i++) {
        if (abc) {
            const float v[9] = {T[i].a.x, T[i].a.y, T[i].a.z, T[i].b.x, T[i].b.y, T[i].b.z, T[i].c.x, T[i].c.y, T[i].c.z};
            std::memcpy(abc + 9 * i, v, sizeof v);
        }
        if (color) color[3 * i] = T[i].color.x, color[3 * i + 1] = T[i].color.y, color[3 * i + 2] = T[i].color.z;
        if (emit) emit[i] = T[i].isEmitting ? 1 : 0;
    }
}

void ph_scene_get_derived(const ph_scene *cs, float *ab, float *ac, float *normal) {
    ph_scene *s = const_cast<ph_scene *>(cs);
    s->ensureDerived();
    const auto &T = s->s.triangles;
    for (size_t i = 0; i < T.size(); i++) {
        if (ab) ab[3 * i] = T[i].ab.x, ab[3 * i + 1] = T[i].ab.y, ab[3 * i + 2] = T[i].ab.z;
        if (ac) ac[3 * i] = T[i].ac.x, ac[3 * i + 1] = T[i].ac.y, ac[3 * i + 2] = T[i].ac.z;
        if (normal) normal[3 * i] = T[i].normal.x, normal[3 * i + 1] = T[i].normal.y, normal[3 * i + 2] = T[i].normal.z;
    }
}

void ph_scene_info(const ph_scene *s, double *o) {
    const pray::Scene &S = s->s;
    o[0] = S.mode;
    o[1] = (double)S.width;
    o[2] = (double)S.height;
    o[3] = (double)S.field_of_view;
    o[4] = (double)S.maxDepth;
    o[5] = (double)S.numSamples;
    o[6] = (double)S.lights.size();
    o[7] = S.backgroundColor.x, o[8] = S.backgroundColor.y, o[9] = S.backgroundColor.z;
    o[10] = S.detEps;
    o[11] = S.rrDepth == SIZE_MAX ? -1.0 : (double)S.rrDepth;
    o[12] = (double)S.seed;
    o[13] = S.treeKind;
    o[14] = o[15] = 0;
}

void ph_scene_set_render(ph_scene *s, uint32_t width, uint32_t height, uint32_t spp, uint32_t max_depth, int rr_depth) {
    if (width) s->s.width = width;
    if (height) s->s.height = height;
    if (spp) s->s.numSamples = spp;
    if (max_depth) s->s.maxDepth = max_depth;
    if (rr_depth >= 0) s->s.rrDepth = (size_t)rr_depth;
}

void ph_scene_get_lights(const ph_scene *s, float *l) {
    for (size_t i = 0; i < s->s.lights.size(); i++) {
        const auto &L = s->s.lights[i];
        const float v[6] = {L.position.x, L.position.y, L.position.z, L.color.x, L.color.y, L.color.z};
        std::memcpy(l + 6 * i, v, sizeof v);
    }
}

void ph_camera(const ph_scene *s, float *cam16) { pray::cameraParams(s->s, cam16); }

void ph_prim_rays(const ph_scene *s, float *rays) {
    std::vector<float> r;
    pray::calculatePrimRays(s->s, r);
    std::memcpy(rays, r.data(), r.size() * sizeof(float));
}

ph_tree *ph_build(ph_scene *s, int kind) { return ph_build_depth(s, kind, (int)pray::KDTree::MAX_DEPTH); }

ph_tree *ph_build_depth(ph_scene *s, int kind, int max_depth) { return ph_build_ex(s, kind, max_depth, 0, 0.f, 0.f); }

ph_tree *ph_build_ex(ph_scene *s, int kind, int max_depth, int leaf_size, float traversal_cost,
                     float intersection_cost) {
    s->ensureDerived();
    auto *t = new ph_tree();
    pray::Scene &S = s->s;
    const int saved = S.treeKind;
    if (kind >= 0) S.treeKind = kind;
    t->kind = pray::chooseTree(S);
    S.treeKind = saved;
    if (t->kind == 2)
        t->tree = std::make_unique<pray::KDTreeBinnedSAH>();
    else if (t->kind == 1)
        t->tree = std::make_unique<pray::KDTreeSAH>();
    else
        t->tree = std::make_unique<pray::KDTree>();
    t->tree->depthLimit = max_depth > 0 ? (size_t)max_depth : pray::KDTree::MAX_DEPTH;
    if (leaf_size > 0) t->tree->leafSize = (size_t)leaf_size;
    if (auto *sah = dynamic_cast<pray::KDTreeSAH *>(t->tree.get())) {
        if (traversal_cost > 0) sah->traversalCost = traversal_cost;
        if (intersection_cost > 0) sah->intersectionCost = intersection_cost;
    }
    if (auto *bin = dynamic_cast<pray::KDTreeBinnedSAH *>(t->tree.get())) {
        if (traversal_cost > 0) bin->traversalCost = traversal_cost;
        if (intersection_cost > 0) bin->intersectionCost = intersection_cost;
    }
    t->tree->build(S.triangles);
    t->flat = pray::buildGPUTreeFromKDTree(*t->tree);
    return t;
}

void ph_tree_free(ph_tree *t) { delete t; }

void ph_tree_stats(const ph_tree *t, uint64_t *o) {
    o[0] = (uint64_t)t->kind;
    if (t->tree) {
        o[1] = t->tree->nnodes;
        o[2] = t->tree->maxdepth;
        o[3] = t->tree->leafNodes;
        o[4] = t->tree->emptyLeafNodes;
        o[5] = t->tree->triangleCount;
    } else {  // loaded from a file: the flat layout's own counts
        o[1] = 2 * (uint64_t)t->flat.numLeaves - 1;
        o[2] = t->flat.maxDepth;
        o[3] = t->flat.numLeaves;
        o[4] = t->flat.numEmptyLeaves;
        o[5] = t->flat.numTris();
    }
    o[6] = t->flat.numNodes();
    o[7] = t->flat.leafRefs.size();
}

int ph_tree_save(const ph_tree *t, const char *path, char *err, uint64_t err_len) {
    std::string e;
    if (!t || !path) return -1;
    if (pray::saveFlat(t->flat, t->kind, path, e)) return 0;
    if (err && err_len) std::snprintf(err, err_len, "%s", e.c_str());
    return -1;
}

ph_tree *ph_tree_load(const char *path, char *err, uint64_t err_len) {
    auto *t = new ph_tree();
    std::string e;
    if (!path || !pray::loadFlat(path, t->flat, t->kind, e)) {
        if (err && err_len) std::snprintf(err, err_len, "%s", path ? e.c_str() : "null path");
        delete t;
        return nullptr;
    }
    return t;
}

uint64_t ph_scene_fingerprint(const ph_scene *s) { return s ? pray::renderFingerprint(s->s) : 0; }

int ph_write_pfm(const char *path, uint32_t w, uint32_t h, const float *rgb) {
    if (!path || !rgb || !w || !h) return -1;
    return pray::writePFM(path, w, h, rgb) ? 0 : -1;
}

int ph_read_pfm(const char *path, uint32_t *wh, float *rgb, char *err, uint64_t err_len) {
    std::string e;
    uint32_t w = 0, h = 0;
    std::vector<float> v;
    if (!path || !wh || !pray::readPFM(path, w, h, v, e)) {
        if (err && err_len) std::snprintf(err, err_len, "%s", path && wh ? e.c_str() : "null argument");
        return -1;
    }
    wh[0] = w, wh[1] = h;
    if (rgb) std::memcpy(rgb, v.data(), v.size() * 4);
    return 0;
}

int ph_accum_save(const char *path, uint32_t w, uint32_t h, uint32_t samples, uint64_t fingerprint, const float *sum,
                  char *err, uint64_t err_len) {
    std::string e;
    if (!path || (!sum && (uint64_t)w * h != 0)) {
        if (err && err_len) std::snprintf(err, err_len, "null argument");
        return -1;
    }
    if (pray::saveAccum(path, w, h, samples, fingerprint, sum, e)) return 0;
    if (err && err_len) std::snprintf(err, err_len, "%s", e.c_str());
    return -1;
}

int ph_accum_load(const char *path, uint32_t *info3, uint64_t *fingerprint, float *sum, char *err, uint64_t err_len) {
    std::string e;
    uint32_t w = 0, h = 0, n = 0;
    uint64_t fp = 0;
    std::vector<float> v;
    if (!path || !info3 || !pray::loadAccum(path, w, h, n, fp, v, e)) {
        if (err && err_len) std::snprintf(err, err_len, "%s", path && info3 ? e.c_str() : "null argument");
        return -1;
    }
    info3[0] = w, info3[1] = h, info3[2] = n;
    if (fingerprint) *fingerprint = fp;
    if (sum) std::memcpy(sum, v.data(), v.size() * 4);
    return 0;
}

void ph_tree_export(const ph_tree *t, uint64_t *sizes, int32_t *ni, float *nf, uint32_t *tris) {
    pray::TreeExport ex;
    if (t->tree) pray::exportTree(t->tree->root.get(), ex);
    const size_t n = ex.is_leaf.size();
    if (sizes) sizes[0] = n, sizes[1] = ex.tris.size();
    if (ni)
        for (size_t i = 0; i < n; i++) {
            ni[4 * i] = ex.is_leaf[i];
            ni[4 * i + 1] = ex.axis[i];
            ni[4 * i + 2] = (int32_t)ex.first[i];
            ni[4 * i + 3] = (int32_t)ex.count[i];
        }
    if (nf)
        for (size_t i = 0; i < n; i++) {
            nf[7 * i] = ex.split[i];
            for (int k = 0; k < 6; k++) nf[7 * i + 1 + k] = ex.box[6 * i + k];
        }
    if (tris) std::memcpy(tris, ex.tris.data(), ex.tris.size() * 4);
}

void ph_flat_get(const ph_tree *t, uint64_t *sizes, uint32_t *nodes, uint32_t *refs, float *a, float *ab, float *ac,
                 float *normal, float *color, uint8_t *emit, float *root_box) {
    const pray::KDTreeGPU &g = t->flat;
    if (sizes) sizes[0] = g.numNodes(), sizes[1] = g.leafRefs.size(), sizes[2] = g.numTris(), sizes[3] = g.maxDepth;
    auto cp = [](void *dst, const void *src, size_t bytes) {
        if (dst && bytes) std::memcpy(dst, src, bytes);
    };
    cp(nodes, g.nodes.data(), g.nodes.size() * 4);
    cp(refs, g.leafRefs.data(), g.leafRefs.size() * 4);
    cp(a, g.triA.data(), g.triA.size() * 4);
    cp(ab, g.triAB.data(), g.triAB.size() * 4);
    cp(ac, g.triAC.data(), g.triAC.size() * 4);
    cp(normal, g.triNormal.data(), g.triNormal.size() * 4);
    cp(color, g.triColor.data(), g.triColor.size() * 4);
    cp(emit, g.triEmit.data(), g.triEmit.size());
    cp(root_box, g.rootBox, sizeof g.rootBox);
}

int ph_upload(const ph_tree *t, pray_ctx *ctx, int with_shading, int *scene_id) {
    const pray::KDTreeGPU &g = t->flat;
    if (with_shading && g.triNormal.size() != 3 * (size_t)g.numTris()) with_shading = 0;  // file without shading
    return pray_upload_scene(ctx, g.nodes.data(), g.numNodes(), g.leafRefs.data(), (uint32_t)g.leafRefs.size(),
                             g.triA.data(), g.triAB.data(), g.triAC.data(), with_shading ? g.triNormal.data() : nullptr,
                             with_shading ? g.triColor.data() : nullptr, with_shading ? g.triEmit.data() : nullptr,
                             g.numTris(), g.rootBox, g.maxDepth, scene_id);
}

}  // extern "C"
