// Pray CLI: `Pray scene.json out.bmp` (reference src/main.cpp:23-109), plus
// `Pray --gen <terrain|spheres|cornell|random> <seed> <out.json> [n]` to write
// the procedural scenes as JSON + OBJ + MTL.
//
// Path-tracing scenes take three more options after the two file names:
//   --hdr out.pfm     also write the linear mean radiance (before the clamp)
//   --accum state     progressive rendering: continue the running sum kept
//                     in `state` (created when absent) and save it again
//   --spp N           samples to add in this run (default: the scene's)
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

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
    std::string hdrPath, accumPath;
    long long sppOverride = 0;
    bool badArgs = argc < 3;
    for (int i = 3; i < argc && !badArgs; i++) {
        const std::string a = argv[i];
        if (i + 1 >= argc) {
            badArgs = true;
        } else if (a == "--hdr") {
            hdrPath = argv[++i];
        } else if (a == "--accum") {
            accumPath = argv[++i];
        } else if (a == "--spp") {
            sppOverride = std::strtoll(argv[++i], nullptr, 10);
            badArgs = sppOverride <= 0 || sppOverride >= (1ll << 24);
        } else {
            badArgs = true;
        }
    }
    if (badArgs) {
        std::cout << "usage: " << argv[0] << " <input.json> <output.bmp> [--hdr out.pfm] [--accum state] [--spp N]\n";
        return 1;
    }
    const bool progressive = !hdrPath.empty() || !accumPath.empty() || sppOverride > 0;
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
    if (progressive && scene.mode != pray::PATHTRACING) {
        std::cerr << "--hdr, --accum and --spp apply to path-tracing scenes\n";
        return 1;
    }
    if (sppOverride > 0) scene.numSamples = (size_t)sppOverride;
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
            // The reference falls back to its CPU tracer here
            // (scene_preprocessor.cpp:42-49); this drop-in has none.
            std::cerr << "the GPU cannot take this scene (no usable GPU, or not enough free device memory)\n";
            pray::shutdownGPU();
            return 2;
        }
        t1 = clk::now();
        std::cout << "Preprocessing time: " << ms(t0, t1) << " ms" << std::endl;

        Image image(3 * scene.width * scene.height);
        std::cout << "\nRendering scene with " << (scene.mode == pray::RAYTRACING ? "ray" : "path")
                  << " tracing on the GPU..." << std::endl;
        t0 = clk::now();
        std::vector<float> hdr;
        if (progressive) {
            // Resume from the saved running sum when it belongs to this scene.
            std::vector<float> accum;
            uint32_t first = 0;
            const uint64_t fingerprint = pray::renderFingerprint(scene);
            if (!accumPath.empty()) {
                if (FILE *probe = std::fopen(accumPath.c_str(), "rb")) {
                    std::fclose(probe);
                    uint32_t w = 0, h = 0;
                    uint64_t fp = 0;
                    if (!pray::loadAccum(accumPath, w, h, first, fp, accum, err)) throw std::runtime_error(err);
                    if (w != scene.width || h != scene.height || fp != fingerprint)
                        throw std::runtime_error(accumPath + " belongs to another scene or render setting");
                    if ((uint64_t)first + scene.numSamples >= (1ull << 24))
                        throw std::runtime_error("more than 2^24 samples in total");
                    std::cout << "Resuming after " << first << " samples" << std::endl;
                }
            }
            pray::pathTraceProgressive(image, scene, first, (uint32_t)scene.numSamples, accum,
                                       hdrPath.empty() ? nullptr : &hdr);
            const uint32_t total = first + (uint32_t)scene.numSamples;
            if (!accumPath.empty() &&
                !pray::saveAccum(accumPath, (uint32_t)scene.width, (uint32_t)scene.height, total, fingerprint,
                                 accum.data(), err))
                throw std::runtime_error(err);
            std::cout << "Accumulated " << total << " samples" << std::endl;
        } else {
            pray::render(image, scene);
        }
        t1 = clk::now();
        std::cout << "Finished rendering. (" << ms(t0, t1) << " ms)" << std::endl;

        std::cout << "Writing result image..." << std::endl;
        t0 = clk::now();
        if (!pray::writeBMP(argv[2], (uint32_t)scene.width, (uint32_t)scene.height, image.data())) {
            std::cerr << "writing " << argv[2] << " failed\n";
            return 1;
        }
        if (!hdrPath.empty() && !pray::writePFM(hdrPath, (uint32_t)scene.width, (uint32_t)scene.height, hdr.data())) {
            std::cerr << "writing " << hdrPath << " failed\n";
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
