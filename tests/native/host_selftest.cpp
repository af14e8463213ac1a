// Native self-test of the host library's CPU parts and the oracle, built
// under sanitizers by tests/native/Makefile (AddressSanitizer +
// UndefinedBehaviorSanitizer with g++, ThreadSanitizer with clang++ and the
// LLVM OpenMP runtime, whose Archer tool tells ThreadSanitizer about OpenMP
// task and barrier synchronisation).  It exercises what runs on the CPU in
// the drop-in: procedural scenes and the triangle precompute
// (scene_preprocessor.cpp:74-96), the median / SAH / binned-SAH tree builders
// with their OpenMP tasks (kd_tree.cpp:25-84, kd_tree_sah.cpp:32-296), the
// flat device layout, the flat tree file, the scene loader on the example
// scenes, PFM and accumulation files; and the oracle's traversal twin, brute
// force and GPU-build twins.  Results are checked, not only run: every flat
// tree's closest hits equal brute force, and builds with 1 and with many
// threads give identical layouts.
//
// usage: host_selftest <golden scenes directory> <scratch directory>
#include <omp.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../nostrum_pray_amd/host/gpu_layout.hpp"
#include "../../nostrum_pray_amd/host/kd_tree.hpp"
#include "../../nostrum_pray_amd/host/scene.hpp"

extern "C" {
void orc_gen_rays(uint64_t seed, const float *box6, uint64_t n, float *rays);
void orc_bf_closest(const float *A, const float *AB, const float *AC, uint64_t n_tris, const float *rays, uint64_t n,
                    float eps, int32_t *out_idx, float *out_t);
void orc_flat_trace(const uint32_t *nodes, const uint32_t *refs, const float *A, const float *AB, const float *AC,
                    const float *box6, uint64_t n_tris, const float *rays, uint64_t n, int mode, float eps,
                    const float *lim2, int32_t *out_idx, float *out_t, uint32_t *counts);
uint32_t orc_lbvh_build(const float *abc, uint64_t n, uint32_t leaf_size, uint32_t *nodes, uint32_t *refs,
                        float *box);
uint64_t orc_sah_build(const float *abc, uint64_t n, uint32_t leaf_size, uint32_t max_depth, float ct, float ci,
                       uint32_t *nodes_out, uint64_t nodes_cap, uint32_t *refs_out, uint64_t refs_cap,
                       uint64_t *n_refs, uint32_t *depth, float *box);
}

namespace {

int failures = 0;

void check(bool ok, const std::string &what) {
    if (!ok) {
        std::fprintf(stderr, "FAIL: %s\n", what.c_str());
        ++failures;
    }
}

std::unique_ptr<pray::KDTree> make_tree(int kind, size_t depth, size_t leaf, float ct) {
    std::unique_ptr<pray::KDTree> t;
    if (kind == 2) {
        auto b = std::make_unique<pray::KDTreeBinnedSAH>();
        b->traversalCost = ct;
        t = std::move(b);
    } else if (kind == 1) {
        t = std::make_unique<pray::KDTreeSAH>();
    } else {
        t = std::make_unique<pray::KDTree>();
    }
    t->depthLimit = depth;
    t->leafSize = leaf;
    return t;
}

// Per-triangle a, ab, ac as the flat layout stores them.
void tri_arrays(const pray::KDTreeGPU &g, std::vector<float> &a, std::vector<float> &ab, std::vector<float> &ac) {
    a = g.triA, ab = g.triAB, ac = g.triAC;
}

// Closest hits of the flat tree's traversal twin against brute force on
// incoherent rays through the root box.
void check_hits(const pray::KDTreeGPU &g, const std::string &name) {
    const uint64_t n = 1500;
    float box6[6];
    for (int k = 0; k < 3; k++) {
        const float ext = g.rootBox[k + 3] - g.rootBox[k];
        box6[k] = g.rootBox[k] - 0.05f * ext;
        box6[k + 3] = 1.1f * ext;
    }
    std::vector<float> rays(6 * n);
    orc_gen_rays(7, box6, n, rays.data());
    std::vector<int32_t> fi(n), bi(n);
    std::vector<float> ft(n), bt(n);
    std::vector<uint32_t> cnt(2 * n);
    std::vector<float> a, ab, ac;
    tri_arrays(g, a, ab, ac);
    orc_flat_trace(g.nodes.data(), g.leafRefs.data(), a.data(), ab.data(), ac.data(), g.rootBox, g.numTris(),
                   rays.data(), n, 0, 1e-9f, nullptr, fi.data(), ft.data(), cnt.data());
    orc_bf_closest(a.data(), ab.data(), ac.data(), g.numTris(), rays.data(), n, 1e-9f, bi.data(), bt.data());
    uint64_t hits = 0;
    bool same = true;
    for (uint64_t i = 0; i < n; i++) {
        same = same && fi[i] == bi[i] && (fi[i] < 0 || ft[i] == bt[i]);
        hits += fi[i] >= 0;
    }
    check(same, name + ": flat traversal equals brute force");
    check(hits > 0, name + ": some rays hit");
}

bool same_layout(const pray::KDTreeGPU &x, const pray::KDTreeGPU &y) {
    return x.nodes == y.nodes && x.leafRefs == y.leafRefs && x.triA == y.triA && x.maxDepth == y.maxDepth;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <golden scenes dir> <scratch dir>\n", argv[0]);
        return 2;
    }
    const std::string scenes = argv[1], tmp = argv[2];
    const int threads = omp_get_max_threads();
    std::printf("threads %d\n", threads);

    // ---- procedural scenes x builders, single- and multi-threaded builds
    struct Case {
        const char *name;
        int kind;
        size_t depth, leaf;
        float ct;
    };
    const Case cases[] = {{"median", 0, 20, 32, 0.f},      {"sah", 1, 20, 32, 0.f},
                          {"binned", 2, 40, 8, 2.f},       {"binned-leaf1", 2, 40, 1, 0.5f},
                          {"median-deep", 0, 30, 4, 0.f}};
    for (int s = 0; s < 4; s++) {
        pray::Scene sc;
        const char *sname = s == 0 ? "spheres" : s == 1 ? "terrain64" : s == 2 ? "cornell" : "random";
        if (s == 0) pray::makeSpheresScene(sc, 1);
        if (s == 1) pray::makeTerrainScene(sc, 64, 9);
        if (s == 2) pray::makeCornellScene(sc, 2, 6000);
        if (s == 3) pray::makeRandomScene(sc, 4, 5000);
        pray::initialiseTriangles(sc);
        for (const Case &c : cases) {
            const std::string name = std::string(sname) + "/" + c.name;
            omp_set_num_threads(1);
            auto t1 = make_tree(c.kind, c.depth, c.leaf, c.ct);
            t1->build(sc.triangles);
            const pray::KDTreeGPU g1 = pray::buildGPUTreeFromKDTree(*t1);
            omp_set_num_threads(threads);
            auto tn = make_tree(c.kind, c.depth, c.leaf, c.ct);
            tn->build(sc.triangles);
            const pray::KDTreeGPU gn = pray::buildGPUTreeFromKDTree(*tn);
            check(same_layout(g1, gn), name + ": 1-thread and many-thread builds give the same layout");
            check_hits(gn, name);
            pray::TreeExport ex;
            pray::exportTree(tn->root.get(), ex);
            check(!ex.is_leaf.empty(), name + ": export");
        }
        // the oracle's twins of the GPU builds
        std::vector<float> abc(9 * sc.triangles.size());
        for (size_t i = 0; i < sc.triangles.size(); i++) {
            const pray::Triangle &t = sc.triangles[i];
            const float v[9] = {t.a.x, t.a.y, t.a.z, t.b.x, t.b.y, t.b.z, t.c.x, t.c.y, t.c.z};
            std::memcpy(&abc[9 * i], v, sizeof v);
        }
        const uint64_t n = sc.triangles.size();
        std::vector<uint32_t> ln(8 * n + 64), lr(n + 2);
        float lbox[6];
        const uint32_t ldepth = orc_lbvh_build(abc.data(), n, 8, ln.data(), lr.data(), lbox);
        check(ldepth > 0 && ldepth < 200, std::string(sname) + ": Morton twin");
        const uint64_t cap = 64 * n + 1024;
        std::vector<uint32_t> sn(cap), sr(cap);
        uint64_t nrefs = 0;
        uint32_t sdepth = 0;
        float sbox[6];
        const uint64_t nn = orc_sah_build(abc.data(), n, 8, 40, 2.f, 1.f, sn.data(), cap, sr.data(), cap, &nrefs,
                                          &sdepth, sbox);
        check(nn > 1 && nrefs >= n && sdepth <= 40, std::string(sname) + ": SAH twin");
    }

    // ---- flat tree file round trip
    {
        pray::Scene sc;
        pray::makeTerrainScene(sc, 48, 3);
        pray::initialiseTriangles(sc);
        auto t = make_tree(2, 40, 8, 2.f);
        t->build(sc.triangles);
        const pray::KDTreeGPU g = pray::buildGPUTreeFromKDTree(*t);
        std::string err;
        const std::string path = tmp + "/selftest.pflat";
        check(pray::saveFlat(g, 2, path, err), "flat file save: " + err);
        pray::KDTreeGPU back;
        int kind = -1;
        check(pray::loadFlat(path, back, kind, err), "flat file load: " + err);
        check(kind == 2 && same_layout(g, back) && back.triColor == g.triColor, "flat file round trip");
        check(!pray::loadFlat(tmp + "/missing.pflat", back, kind, err), "missing flat file refused");
    }

    // ---- the example scenes through the loader and the reference policy
    for (const char *f : {"head_fov30.json", "cube_200x150.json", "plant_256.json", "pt_cornell.json"}) {
        pray::Scene sc;
        std::string err;
        const bool ok = pray::readSceneDescription(scenes + "/" + f, sc, err);
        check(ok && !sc.triangles.empty(), std::string("scene ") + f + ": " + err);
        if (!ok) continue;
        pray::preprocessScene(sc, false);
        check(sc.triangleDB != nullptr, std::string("scene ") + f + ": tree");
        if (sc.triangleDB) check_hits(pray::buildGPUTreeFromKDTree(*sc.triangleDB), f);
        std::vector<float> rays;
        pray::calculatePrimRays(sc, rays);
        check(rays.size() == 6 * sc.width * sc.height, std::string("scene ") + f + ": primary rays");
    }

    // ---- PFM and accumulation files
    {
        const uint32_t w = 7, h = 5;
        std::vector<float> img(3 * w * h);
        for (size_t i = 0; i < img.size(); i++) img[i] = 0.25f * (float)i - 3.f;
        std::string err;
        check(pray::writePFM(tmp + "/selftest.pfm", w, h, img.data()), "PFM write");
        uint32_t w2 = 0, h2 = 0;
        std::vector<float> back;
        check(pray::readPFM(tmp + "/selftest.pfm", w2, h2, back, err) && w2 == w && h2 == h && back == img,
              "PFM round trip: " + err);
        check(pray::saveAccum(tmp + "/selftest.pacc", w, h, 9, 0x1234abcdull, img.data(), err), "accum save");
        uint32_t s = 0;
        uint64_t fp = 0;
        back.clear();
        check(pray::loadAccum(tmp + "/selftest.pacc", w2, h2, s, fp, back, err) && s == 9 && fp == 0x1234abcdull &&
                  back == img,
              "accum round trip: " + err);
    }

    std::printf("%s (%d failures)\n", failures ? "FAILED" : "selftest ok", failures);
    return failures ? 1 : 0;
}
