// Launch of the megakernel frame (settings.use_wavefront_pathtracing = 0).
#pragma once

#include "stages.h"

namespace wf {

struct MegaLaunch {
    TraceScene scene;
    uint2* spill;              // global overflow stack (level-major)
    uint32_t spill_per_thread;
    uint32_t stack_cap;
    uint32_t* work;            // [0] next path, [1] stack overflows
    uint32_t visits;           // interior visits per traversal pass
    uint32_t refill_min;       // idle lanes of a wave that trigger a refill
    uint32_t grid;             // persistent blocks
};

hipError_t mega_occupancy(const ExtendConfig& cfg, int* blocks_per_cu);
hipError_t launch_megakernel(const FrameArgs& A, const MegaLaunch& M, const ExtendConfig& cfg, hipStream_t stream);

} // namespace wf
