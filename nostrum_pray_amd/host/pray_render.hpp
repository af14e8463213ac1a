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
void shutdownGPU();
}  // namespace pray
