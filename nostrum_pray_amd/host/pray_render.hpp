// Renderer entry points of the drop-in host (reference include/pray.h:50-53,
// include/path_trace.cu.h:11-12).
#pragma once

#include <cstdint>
#include <vector>

#include "scene.hpp"

using Image = std::vector<uint8_t>;

namespace pray {
void render(Image &image, const Scene &scene);
void rayTrace(Image &image, const Scene &scene);
void pathTrace(Image &image, const Scene &scene);
bool isComputableOnGPU(const Scene &scene);
void pathTraceOnGPURecursive(Image &image, const Scene &scene);
// Path tracing of samples first .. first+spp-1 (no reference counterpart):
// accum holds the running radiance sum (3 floats per pixel; resized and
// zeroed when first == 0) and is updated in place; image and hdr (optional)
// receive the clamped 8-bit and the linear mean over all first+spp samples.
void pathTraceProgressive(Image &image, const Scene &scene, uint32_t first, uint32_t spp, std::vector<float> &accum,
                          std::vector<float> *hdr);
void shutdownGPU();
}  // namespace pray
