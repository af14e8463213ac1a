// Renderer entry points of the drop-in host (reference include/pray.h:50-53,
// include/path_trace.cu.h:11-12), routed through libpray_hip.  There is no
// CPU tracing path: when the GPU cannot take a scene the call fails with a
// message instead of silently falling back.
#include <cstdio>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pray_hip.h"
#include "gpu_layout.hpp"
#include "pray_render.hpp"

namespace pray {

namespace {
pray_ctx *g_ctx = nullptr;

pray_ctx *ctx() {
    if (!g_ctx) {
        const int rc = pray_init(0, &g_ctx);
        if (rc != PRAY_OK) throw std::runtime_error("GPU initialisation failed (code " + std::to_string(rc) + ")");
    }
    return g_ctx;
}

void check(int rc, const char *what) {
    if (rc != PRAY_OK) throw std::runtime_error(std::string(what) + ": " + pray_last_error(g_ctx));
}

int upload(const Scene &scene, const KDTreeGPU &g) {
    int id = -1;
    check(pray_upload_scene(ctx(), g.nodes.data(), g.numNodes(), g.leafRefs.data(), (uint32_t)g.leafRefs.size(),
                            g.triA.data(), g.triAB.data(), g.triAC.data(), g.triNormal.data(), g.triColor.data(),
                            g.triEmit.data(), g.numTris(), g.rootBox, g.maxDepth, &id),
          "scene upload");
    return id;
}
}  // namespace

// The reference's eligibility check (path_trace.cu:557-583): a bool the
// caller acts on (scene_preprocessor.cpp:42-49), false when there is no
// usable device or it lacks the memory -- never an exception.
bool isComputableOnGPU(const Scene &scene) {
    if (!g_ctx && pray_init(0, &g_ctx) != PRAY_OK) {
        g_ctx = nullptr;
        return false;
    }
    int ok = 0;
    const uint64_t n = scene.triangles.size();
    if (pray_mem_ok(g_ctx, n, n + n / 2, 4 * (n / 8 + 1), scene.width * scene.height, &ok) != PRAY_OK) return false;
    return ok != 0;
}

void rayTrace(Image &image, const Scene &scene) {
    const KDTreeGPU g = buildGPUTreeFromKDTree(*scene.triangleDB);
    const int id = upload(scene, g);
    float cam[16];
    cameraParams(scene, cam);
    std::vector<float> lights;
    for (const LightSource &l : scene.lights) {
        const float v[6] = {l.position.x, l.position.y, l.position.z, l.color.x, l.color.y, l.color.z};
        lights.insert(lights.end(), v, v + 6);
    }
    const float bg[3] = {scene.backgroundColor.x, scene.backgroundColor.y, scene.backgroundColor.z};
    image.assign(scene.width * scene.height * 3, 0);
    check(pray_render_rt(ctx(), id, cam, (uint32_t)scene.width, (uint32_t)scene.height, lights.data(),
                         (uint32_t)scene.lights.size(), bg, scene.detEps, image.data(), nullptr),
          "ray tracing");
    pray_free_scene(ctx(), id);
}

void pathTraceOnGPURecursive(Image &image, const Scene &scene) {
    const KDTreeGPU g = buildGPUTreeFromKDTree(*scene.triangleDB);
    const int id = upload(scene, g);
    float cam[16];
    cameraParams(scene, cam);
    const float bg[3] = {scene.backgroundColor.x, scene.backgroundColor.y, scene.backgroundColor.z};
    image.assign(scene.width * scene.height * 3, 0);
    const uint32_t rr = scene.rrDepth == SIZE_MAX ? 0xFFFFFFFFu : (uint32_t)scene.rrDepth;
    check(pray_render_pt(ctx(), id, cam, (uint32_t)scene.width, (uint32_t)scene.height, (uint32_t)scene.numSamples,
                         (uint32_t)scene.maxDepth, rr, scene.seed, bg, scene.detEps, nullptr, image.data()),
          "path tracing");
    pray_free_scene(ctx(), id);
}

void pathTrace(Image &image, const Scene &scene) { pathTraceOnGPURecursive(image, scene); }

void pathTraceProgressive(Image &image, const Scene &scene, uint32_t first, uint32_t spp, std::vector<float> &accum,
                          std::vector<float> *hdr) {
    const size_t n3 = 3 * scene.width * scene.height;
    image.assign(n3, 0);
    if (hdr) hdr->assign(n3, 0.f);
    if (first == 0) accum.assign(n3, 0.f);
    if (accum.size() != n3) throw std::runtime_error("accumulation buffer does not match the image size");
    if (!scene.triangleDB || scene.triangleDB->isEmpty()) {
        // Every sample of every pixel is the background: the same fp32 sums
        // the GPU would form, per channel.
        const float bg[3] = {scene.backgroundColor.x, scene.backgroundColor.y, scene.backgroundColor.z};
        for (int ch = 0; ch < 3; ch++) {
            float acc = accum[ch];
            for (uint32_t k = 0; k < spp; k++) acc = acc + bg[ch];
            const float m = acc / (float)(first + spp);
            const float lo = m < 1.f ? m : 1.f;
            const float cl = 0.f < lo ? lo : 0.f;
            for (size_t i = ch; i < n3; i += 3) {
                accum[i] = acc;
                image[i] = static_cast<uint8_t>(cl * 255.f + 0.5f);
                if (hdr) (*hdr)[i] = m;
            }
        }
        return;
    }
    const KDTreeGPU g = buildGPUTreeFromKDTree(*scene.triangleDB);
    const int id = upload(scene, g);
    float cam[16];
    cameraParams(scene, cam);
    const float bg[3] = {scene.backgroundColor.x, scene.backgroundColor.y, scene.backgroundColor.z};
    const uint32_t rr = scene.rrDepth == SIZE_MAX ? 0xFFFFFFFFu : (uint32_t)scene.rrDepth;
    const int rc = pray_render_pt_progressive(ctx(), id, cam, (uint32_t)scene.width, (uint32_t)scene.height, first,
                                              spp, (uint32_t)scene.maxDepth, rr, scene.seed, bg, scene.detEps,
                                              accum.data(), hdr ? hdr->data() : nullptr, image.data());
    pray_free_scene(ctx(), id);
    check(rc, "path tracing");
}

// render (pray.cpp:617-637): an empty scene is filled with the rounded
// background colour.
void render(Image &image, const Scene &scene) {
    if (!scene.triangleDB || scene.triangleDB->isEmpty()) {
        const Vec3 c = scene.backgroundColor * 255.f + 0.5f;
        image.assign(scene.width * scene.height * 3, 0);
        for (size_t i = 0; i < scene.width * scene.height; i++) {
            image[3 * i] = static_cast<uint8_t>(c.x);
            image[3 * i + 1] = static_cast<uint8_t>(c.y);
            image[3 * i + 2] = static_cast<uint8_t>(c.z);
        }
        return;
    }
    if (scene.mode == PATHTRACING)
        pathTrace(image, scene);
    else
        rayTrace(image, scene);
}

void shutdownGPU() {
    if (g_ctx) pray_destroy(g_ctx);
    g_ctx = nullptr;
}

}  // namespace pray
