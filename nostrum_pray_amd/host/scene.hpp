// Scene description and preprocessing of the drop-in host
// (reference include/pray.h:16-48, src/scene_preprocessor.cpp:19-136).
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "geometry.hpp"
#include "kd_tree.hpp"

namespace pray {

typedef enum { RAYTRACING = 0, PATHTRACING = 1 } Mode;
typedef enum { NORMAL = 0, DIFFUSE = 1 } PtMode;

// Thresholds of the reference's scene policy (include/options.h:32-38).
constexpr size_t MAX_DEPTH_DIFFUSE = 1;
constexpr size_t MIN_NUM_TRIANGLES_SAH = 20000;
constexpr size_t MAX_NUM_TRIANGLES_SAH = 1000000;

struct LightSource {
    Vec3 position, color;
    LightSource() = default;
    LightSource(Vec3 p, Vec3 c) : position(p), color(c) {}
};

struct Scene {
    Mode mode = RAYTRACING;
    PtMode ptMode = NORMAL;
    bool useGPU = true;

    size_t maxDepth = 0;
    size_t numSamples = 0;
    uint64_t width = 0, height = 0;
    uint64_t field_of_view = 0;  // horizontal full angle in degrees (integer, as the reference)
    Vec3 camera_position{0.f, 0.f, 0.f}, camera_look{0.f, 0.f, 1.f};
    std::vector<LightSource> lights;
    std::vector<Triangle> triangles;
    std::unique_ptr<KDTree> triangleDB;
    Vec3 backgroundColor{0.f, 0.f, 0.f};

    // Extensions of this build (no reference counterpart): determinant cull
    // used by the GPU tracer, Russian-roulette start depth, RNG seed and an
    // explicit tree choice (-1 = the reference policy).
    float detEps = EPSILON;
    size_t rrDepth = SIZE_MAX;
    uint64_t seed = 0;
    int treeKind = -1;
};

// Per-triangle precompute (scene_preprocessor.cpp:74-96).
void initialiseTriangles(Scene &scene);

// The reference camera (scene_preprocessor.cpp:98-136) reduced to the 16
// floats the GPU generate kernel takes: position, inverse view matrix rows
// 0..2 x cols 0..2, angle, 1/width, 1/height, width/height.
void cameraParams(const Scene &scene, float out16[16]);

// Host restatement of calculatePrimRays: width*height rays (6 floats each).
void calculatePrimRays(const Scene &scene, std::vector<float> &rays);

// Tree choice of the reference (scene_preprocessor.cpp:52-57): median below
// 20k or from 1M triangles (or when the reference GPU path is used), SAH in
// between.  Returns 0 = median, 1 = SAH.
int chooseTree(const Scene &scene);

// initialiseTriangles + tree build (scene_preprocessor.cpp:19-72).  The
// triangles are copied into the tree; the scene keeps its list.
void preprocessScene(Scene &scene, bool verbose = true);

// Procedural scenes.  All are deterministic functions of their arguments;
// triangle order is part of the contract (it fixes tie semantics).
//  terrain   grid x grid cells on [0,1]^2 (x-z), two upward-facing triangles
//            per cell in row-major order, heights = 6-octave value noise
//            normalised to [0, 0.1].
//  spheres   level-4 icosphere r=1 + 16 jittered level-3 icospheres
//            r~U[0.1,0.4] + 2-triangle ground quad (25,602 triangles),
//            per-object albedo U[0.2,0.9], one point light, 512x512, fov 60.
//  cornell   closed room x,z in [-1,1], y in [0,2] (5 walls), two boxes
//            rotated +-18 degrees, emissive ceiling quad (Le = 15); every
//            quad tessellated so that the total is close to target_tris;
//            albedo U[0.1,0.8]; path tracing settings.
//  random    mixes of icospheres, boxes and terrain patches with about
//            n_tris triangles (batched multi-scene workload).
void makeTerrainScene(Scene &scene, uint32_t grid, uint64_t seed);
void makeSpheresScene(Scene &scene, uint64_t seed);
void makeCornellScene(Scene &scene, uint64_t seed, uint32_t target_tris);
void makeRandomScene(Scene &scene, uint64_t seed, uint32_t n_tris);

// Scene I/O (reference src/io.cpp:27-137): JSON description + Wavefront OBJ
// with MTL materials; polygons are fan-triangulated in file order.  Returns
// false with a message on failure instead of exiting.
bool readSceneDescription(const std::string &path, Scene &scene, std::string &err);
// Writes the triangles and materials as OBJ + MTL and a scene JSON next to it.
bool writeSceneFiles(const Scene &scene, const std::string &json_path, std::string &err);

// 24-bit BMP writer (bottom-up rows, BGR), input RGB top row first.
bool writeBMP(const std::string &path, uint32_t w, uint32_t h, const uint8_t *rgb);

// Float RGB images as Portable Float Maps (input / output top row first).
bool writePFM(const std::string &path, uint32_t w, uint32_t h, const float *rgb);
bool readPFM(const std::string &path, uint32_t &w, uint32_t &h, std::vector<float> &rgb, std::string &err);

// Progressive rendering state: the running sum of per-sample radiance of a
// path-traced image after `samples` samples, tagged with the fingerprint of
// everything else that decides the pixels (renderFingerprint), so that a
// resumed render refuses a state from another scene or setting.
uint64_t renderFingerprint(const Scene &scene);
bool saveAccum(const std::string &path, uint32_t w, uint32_t h, uint32_t samples, uint64_t fingerprint,
               const float *sum, std::string &err);
bool loadAccum(const std::string &path, uint32_t &w, uint32_t &h, uint32_t &samples, uint64_t &fingerprint,
               std::vector<float> &sum, std::string &err);

}  // namespace pray
