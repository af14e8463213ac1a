// Pray CLI: `Pray scene.json out.bmp` (reference src/main.cpp:23-109), plus
// `Pray --gen <terrain|spheres|cornell|random> <seed> <out.json> [n]` to write
// the procedural scenes as JSON + OBJ + MTL.
#include <chrono>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>

#include "pray_render.hpp"

using clk = std::chrono::steady_clock;

static long long ms(clk::time_point a, clk::time_point b) {
    return (long long)std::chrono::duration_cast<std::chrono::milliseconds>(b - a).count();
}

int main(int argc, char *argv[]) {
    if (argc >= 5 && std::string(argv[1]) == "--gen") {
        pray::Scene s;
        const std::string kind = argv[2];
        const uint64_t seed = std::strtoull(argv[3], nullptr, 10);
        const uint32_t n = argc >= 6 ? (uint32_t)std::strtoul(argv[5], nullptr, 10) : 0;
        if (kind == "terrain")
            pray::makeTerrainScene(s, n ? n : 64, seed);
        else if (kind == "spheres")
            pray::makeSpheresScene(s, seed);
        else if (kind == "cornell")
            pray::makeCornellScene(s, seed, n ? n : 100000);
        else if (kind == "random")
            pray::makeRandomScene(s, seed, n ? n : 10000);
        else {
            std::cerr << "unknown scene kind " << kind << "\n";
            return 1;
        }
        std::string err;
        if (!pray::writeSceneFiles(s, argv[4], err)) {
            std::cerr << err << "\n";
            return 1;
        }
        return 0;
    }
    if (argc != 3) {
        std::cout << "usage: " << argv[0] << " <input.json> <output.bmp>\n";
        return 1;
    }
    const auto total0 = clk::now();
    std::cout << "--------------------------------------------------------\nINPUT FILE: " << argv[1] << "\n\n";
    std::cout << "Reading scene description..." << std::endl;
    auto t0 = clk::now();
    pray::Scene scene;
    std::string err;
    if (!pray::readSceneDescription(argv[1], scene, err)) {
        std::cerr << err << "\n";
        return 1;
    }
    auto t1 = clk::now();
    std::cout << "Finished reading scene description. (" << ms(t0, t1) << " ms)" << std::endl;
    std::cout << "\nScene information: Resolution: " << scene.width << "x" << scene.height
              << ", # Lights: " << scene.lights.size() << ", # Triangles: " << scene.triangles.size();
    if (scene.mode == pray::PATHTRACING)
        std::cout << ", PATH TRACING: max depth: " << scene.maxDepth << ", # samples: " << scene.numSamples;
    std::cout << "\n" << std::endl;

    try {
        std::cout << "Preprocessing scene..." << std::endl;
        t0 = clk::now();
        pray::preprocessScene(scene);
        if (!pray::isComputableOnGPU(scene)) {
            std::cerr << "the GPU cannot hold this scene\n";
            return 1;
        }
        t1 = clk::now();
        std::cout << "Preprocessing time: " << ms(t0, t1) << " ms" << std::endl;

        Image image(3 * scene.width * scene.height);
        std::cout << "\nRendering scene with " << (scene.mode == pray::RAYTRACING ? "ray" : "path")
                  << " tracing on the GPU..." << std::endl;
        t0 = clk::now();
        pray::render(image, scene);
        t1 = clk::now();
        std::cout << "Finished rendering. (" << ms(t0, t1) << " ms)" << std::endl;

        std::cout << "Writing result image..." << std::endl;
        t0 = clk::now();
        if (!pray::writeBMP(argv[2], (uint32_t)scene.width, (uint32_t)scene.height, image.data())) {
            std::cerr << "writing " << argv[2] << " failed\n";
            return 1;
        }
        t1 = clk::now();
        std::cout << "Finished writing. (" << ms(t0, t1) << " ms)" << std::endl;
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << "\n";
        pray::shutdownGPU();
        return 1;
    }
    pray::shutdownGPU();
    std::cout << "Total time: " << ms(total0, clk::now()) << " ms\n" << std::endl;
    return 0;
}
