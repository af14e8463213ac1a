#include "scene.hpp"

#include <chrono>
#include <iostream>

namespace pray {

void initialiseTriangles(Scene &scene) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < scene.triangles.size(); i++) scene.triangles[i].precompute();
}

namespace {
struct CamSetup {
    Vec3 pos;
    Matrix44 inv;
    float invWidth, invHeight, aspect, angle;
};

// View matrix and constants of scene_preprocessor.cpp:101-121.
CamSetup setupCamera(const Scene &scene) {
    CamSetup c;
    c.pos = scene.camera_position;
    const Vec3 target = scene.camera_position + scene.camera_look;
    const Vec3 up(0, 1, 0);
    Vec3 zaxis = target - scene.camera_position;
    zaxis.normalize();
    Vec3 xaxis = up.cross(zaxis);
    xaxis.normalize();
    const Vec3 yaxis = zaxis.cross(xaxis);
    const Matrix44 view(xaxis.x, yaxis.x, zaxis.x, 0, xaxis.y, yaxis.y, zaxis.y, 0, xaxis.z, yaxis.z, zaxis.z, 0,
                        -xaxis.dot(scene.camera_position), -yaxis.dot(scene.camera_position),
                        -zaxis.dot(scene.camera_position), 1);
    c.inv = view.inverse();
    c.invWidth = 1 / float(scene.width);
    c.invHeight = 1 / float(scene.height);
    c.aspect = scene.width / float(scene.height);
    c.angle = std::tan(pi_f() * 0.5f * scene.field_of_view / 180.f);
    return c;
}
}  // namespace

void cameraParams(const Scene &scene, float o[16]) {
    const CamSetup c = setupCamera(scene);
    o[0] = c.pos.x, o[1] = c.pos.y, o[2] = c.pos.z;
    for (int r = 0; r < 3; r++)
        for (int q = 0; q < 3; q++) o[3 + 3 * r + q] = c.inv.x[r][q];
    o[12] = c.angle, o[13] = c.invWidth, o[14] = c.invHeight, o[15] = c.aspect;
}

void calculatePrimRays(const Scene &scene, std::vector<float> &rays) {
    const CamSetup c = setupCamera(scene);
    rays.resize(scene.width * scene.height * 6);
#pragma omp parallel for schedule(static) collapse(2)
    for (uint64_t yPixel = 0; yPixel < scene.height; yPixel++) {
        for (uint64_t xPixel = 0; xPixel < scene.width; xPixel++) {
            const float x = (2 * ((xPixel + 0.5f) * c.invWidth) - 1) * c.angle;
            const float y = (1 - 2 * ((yPixel + 0.5f) * c.invHeight)) * c.angle / c.aspect;
            const Vec3 dir = c.inv.multDirMatrix(Vec3(x, y, 1));
            const Ray r(c.pos, dir);
            float *o = &rays[6 * (xPixel + yPixel * scene.width)];
            o[0] = r.position.x, o[1] = r.position.y, o[2] = r.position.z;
            o[3] = r.direction.x, o[4] = r.direction.y, o[5] = r.direction.z;
        }
    }
}

int chooseTree(const Scene &scene) {
    if (scene.treeKind >= 0) return scene.treeKind;
    const size_t n = scene.triangles.size();
    return (n < MIN_NUM_TRIANGLES_SAH || n >= MAX_NUM_TRIANGLES_SAH) ? 0 : 1;
}

void preprocessScene(Scene &scene, bool verbose) {
    using clk = std::chrono::steady_clock;
    auto ms = [](clk::time_point a, clk::time_point b) {
        return (long long)std::chrono::duration_cast<std::chrono::milliseconds>(b - a).count();
    };
    auto t0 = clk::now();
    initialiseTriangles(scene);
    auto t1 = clk::now();
    if (verbose) std::cout << "\tInitialising triangles... Done: " << ms(t0, t1) << " ms" << std::endl;

    // Estimator selection (scene_preprocessor.cpp:34-40): NORMAL above a
    // depth limit of 1, DIFFUSE otherwise.  The reference only assigns NORMAL
    // and leaves the member uninitialised in the other case, where it prints
    // that it uses the diffuse estimator; here that case is DIFFUSE, which
    // the path tracer applies at the same depth limits (pray_render_pt).
    scene.ptMode = (scene.mode == PATHTRACING && scene.maxDepth > MAX_DEPTH_DIFFUSE) ? NORMAL : DIFFUSE;

    const int kind = chooseTree(scene);
    if (kind == 1)
        scene.triangleDB = std::make_unique<KDTreeSAH>();
    else
        scene.triangleDB = std::make_unique<KDTree>();
    t0 = clk::now();
    scene.triangleDB->build(scene.triangles);
    t1 = clk::now();
    if (verbose) {
        const KDTree &db = *scene.triangleDB;
        std::cout << "\t\t# nodes: " << db.nnodes << ", tree depth: " << db.maxdepth << ", leaf nodes: " << db.leafNodes
                  << ", non-empty leaf nodes: " << db.leafNodes - db.emptyLeafNodes
                  << ", empty leaf nodes: " << db.emptyLeafNodes << (kind == 1 ? " (SAH)" : " (median)") << std::endl;
        std::cout << "\tDone: " << ms(t0, t1) << " ms" << std::endl;
    }
}

}  // namespace pray
