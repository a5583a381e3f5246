// Arguments of the wavefront stage kernels (one struct per frame).
#pragma once

#include "wf_internal.h"

namespace wf {

// counts[d] = length of the ray queue of bounce d (d <= kMaxBounces);
// counts[kShadowCountBase + d] = length of the shadow-ray queue of bounce d.
constexpr uint32_t kCountSlots = 32;
constexpr uint32_t kShadowCountBase = 16;

struct FrameArgs {
    uint32_t width, height, spp, n_paths;
    wf_view view;
    uint32_t frame_seed, max_bounces, rr_start, nee;
    float rr_survival;
    float env[3];
    uint32_t n_emitters;
    float emitter_area;
    uint32_t accumulate, sample_count;

    float4* thr;                 // per path
    float4* energy;              // per path
    wf_ray* rays[2];             // ping-pong ray queues
    uint32_t* path_of[2];        // queue slot -> path id
    wf_hit* hits;                // extend output for the current queue
    uint32_t* counts;            // kCountSlots
    wf_ray* shadow_rays;
    uint32_t* shadow_path;
    float4* shadow_contrib;
    uint32_t* shadow_visible;

    const GpuShadeInstance* shade_inst;
    const GpuTexture* textures;
    const uint4* emitters;       // (instance, primitive, cdf bits, 0)

    float4* image;               // resolved energy per pixel
    float4* accum;               // progressive accumulator
    uint32_t* final_rgba8;
};

hipError_t launch_clear(const FrameArgs& A, uint32_t grid, hipStream_t s);
hipError_t launch_generate(const FrameArgs& A, uint32_t grid, hipStream_t s);
hipError_t launch_shade(const FrameArgs& A, uint32_t d, uint32_t grid, hipStream_t s);
hipError_t launch_connect_accumulate(const FrameArgs& A, uint32_t d, uint32_t grid, hipStream_t s);
hipError_t launch_resolve_post(const FrameArgs& A, uint32_t grid, hipStream_t s);

} // namespace wf
