// Arguments of the wavefront stage kernels.
//
// A frame renders a contiguous range of scene slots in one set of launches
// (a single scene is the range of length one).  Paths of all scenes share
// the queues; each queue entry carries the TLAS root of its scene, each path
// its scene id.  Everything that changes from frame to frame (seed,
// accumulation weights) lives in device memory, so a captured hipGraph of
// the frame stays valid across frames.
#pragma once

#include "wf_internal.h"

namespace wf {

// counts[d] = length of the ray queue of bounce d (d <= kMaxBounces);
// counts[kShadowCountBase + d] = length of the shadow-ray queue of bounce d;
// counts[kShadeCursorBase + d] = the shade stage's claim cursor over queue d.
constexpr uint32_t kCountSlots = 48;
constexpr uint32_t kShadowCountBase = 16;
constexpr uint32_t kShadeCursorBase = 32;

// one scene slot (device table)
struct alignas(16) GpuScene {
    wf_view view;
    uint32_t width, height, n_pix, pix_off;     // pixels [pix_off, pix_off + n_pix) of the batch images
    uint32_t path_off, n_paths;                 // global path ids [path_off, path_off + n_paths)
    uint32_t tlas_root;                         // global TLAS node index of the scene root
    uint32_t em_off, em_count;                  // emitter records [em_off, em_off + em_count)
    float em_area;                              // total emitter area of the scene
    uint32_t sample_count;                      // accumulation weight 1/sample_count this frame
    uint32_t pad[3];
};

struct FrameDyn {
    uint32_t frame_seed;
    uint32_t pad[3];
};

struct FrameArgs {
    // the scene range of this frame
    const GpuScene* scenes;   // indexed by scene slot
    uint32_t first_scene, n_scenes;
    uint32_t path_base, n_paths;   // global path ids [path_base, path_base + n_paths)
    uint32_t pix_base, n_pix;      // batch pixels [pix_base, pix_base + n_pix)
    uint32_t spp, max_bounces, rr_start, nee;
    uint32_t uniform_diffuse;    // 1: uniform hemisphere sampling (cosine_weighted_diffuse = 0)
    uint32_t view_mode;          // render_view_mode (0 = the path tracer)
    float rr_survival;
    float env[3];                // constant environment (hdr_env_strength x the 1x1 white fallback)
    int32_t env_tex;             // equirect environment texture slot, -1 = the constant
    float env_strength;          // hdr_env_strength
    uint32_t accumulate;
    const FrameDyn* dyn;

    float4* thr;                 // per global path
    float4* energy;              // per global path
    uint32_t* path_scene;        // per global path
    wf_ray* rays[2];             // ping-pong ray queues
    uint32_t* path_of[2];        // queue slot -> global path id
    uint32_t* root_of[2];        // queue slot -> TLAS root node
    wf_hit* hits;                // extend output for the current queue
    uint32_t* counts;            // kCountSlots
    wf_ray* shadow_rays;
    uint32_t* shadow_path;
    uint32_t* shadow_root;
    float4* shadow_contrib;
    uint32_t* shadow_visible;

    const GpuShadeInstance* shade_inst;
    const GpuTexture* textures;
    const uint4* emitters;       // (instance, primitive, cdf bits, 0), per-scene ranges

    float4* image;               // resolved energy per batch pixel
    float4* accum;               // progressive accumulator per batch pixel
    uint32_t* final_rgba8;

    // path-replay adjoint
    float4* lrem;                // per path: estimate still to be collected
    float4* delta;               // per path: dLoss/dpixel / spp
    const float4* dl_dimage;     // per batch pixel
    float4* shadow_unit;         // shadow contribution per unit emission strength
    uint32_t* shadow_light;      // shadow ray's emitter instance
    float4* grad_tex;            // per texel of the differentiated texture
    float* grad_em;              // per instance: dLoss/d emissive_strength
    int32_t grad_tex_slot;       // texture slot differentiated (-1: none)
};

hipError_t launch_clear(const FrameArgs& A, uint32_t grid, hipStream_t s);
hipError_t launch_generate(const FrameArgs& A, uint32_t grid, hipStream_t s);
hipError_t launch_shade(const FrameArgs& A, uint32_t d, uint32_t grid, hipStream_t s);
hipError_t launch_connect_accumulate(const FrameArgs& A, uint32_t d, uint32_t grid, hipStream_t s);
hipError_t launch_resolve_post(const FrameArgs& A, uint32_t grid, hipStream_t s);
hipError_t launch_adjoint_clear(const FrameArgs& A, uint32_t grid, hipStream_t s);
hipError_t launch_adjoint_shade(const FrameArgs& A, uint32_t d, uint32_t grid, hipStream_t s);
hipError_t launch_adjoint_connect(const FrameArgs& A, uint32_t d, uint32_t grid, hipStream_t s);

} // namespace wf
