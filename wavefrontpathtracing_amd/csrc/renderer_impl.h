// The renderer object (the g_renderer state of renderer_common.h:146-243,
// re-expressed over HIP device memory).
#pragma once

#include "wf_internal.h"

#include <cstdarg>
#include <string>
#include <vector>

namespace wf {

constexpr uint32_t kMaxTextures = 65536;   // renderer.cpp:232 texture slotmap size
constexpr uint32_t kStackTotal = 128;      // LDS + spill entries per lane (reference: 64 TLAS + 64 BLAS)
constexpr int kNumStageEvents = 48;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n);
    void release();
};

struct Mesh {
    bool alive = false;
    uint32_t version = 0;
    uint32_t tri_count = 0, node_count = 0, max_leaf = 0;
    float blas_min[3] = {0, 0, 0}, blas_max[3] = {0, 0, 0};
    std::vector<float> pos9;   // host copy of positions (emitter sampling)
    DevBuf nodes, tris, shade;
};

struct Texture {
    bool alive = false;
    uint32_t version = 0;
    uint32_t width = 0, height = 0;
    DevBuf texels;   // RGBA float
};

struct InstanceRec {
    uint32_t mesh = 0;
    float transform[16];
    wf_material_asset material;
    wf_bvh_instance sw;
};

} // namespace wf

struct wf_renderer {
    uint32_t width = 0, height = 0;
    int device = 0;
    int cu_count = 1;
    hipStream_t stream = nullptr;
    hipEvent_t ev[wf::kNumStageEvents] = {};
    wf_render_settings settings{};
    wf_build_options build_opts{8, 0};
    std::string err;

    std::vector<wf::Mesh> meshes;
    std::vector<wf::Texture> textures;
    bool textures_dirty = true;
    std::vector<wf::InstanceRec> instances;

    // scene / frame state
    bool frame_open = false, scene_open = false;
    wf_camera camera{};
    wf_view view{};
    bool view_valid = false;
    uint64_t env_texture = WF_INVALID_HANDLE;
    uint32_t xorshift_seed = 0xC977E154u;   // core/random.h:7
    uint64_t frame_index = 0;
    uint32_t accum_count = 0;
    wf_frame_stats stats{};

    // TLAS + instance records on the device
    wf::TlasData tlas_host;
    wf::DevBuf d_tnodes, d_inst, d_shade_inst, d_textures, d_emitters;
    uint32_t n_emitters = 0;
    float emitter_area = 0.0f;
    bool tlas_valid = false;

    // traversal scratch
    wf::ExtendConfig ext_cfg;
    uint32_t ext_grid = 0;
    uint32_t spill_per_thread = 0;
    wf::DevBuf d_work, d_spill, d_tmp_rays, d_tmp_hits, d_tmp_aux;

    // wavefront buffers (allocated on first render)
    wf::DevBuf d_rays[2], d_paths[2], d_hits, d_thr, d_path_energy, d_counts;
    wf::DevBuf d_shadow_rays, d_shadow_path, d_shadow_contrib, d_shadow_visible;
    wf::DevBuf d_image, d_accum, d_final;
    size_t frame_capacity = 0;
    bool image_valid = false;

    wf_status fail(wf_status s, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
    wf_status hip_fail(hipError_t e, const char* what);
    wf::Mesh* find_mesh(uint64_t h);
    wf::Texture* find_texture(uint64_t h);
    wf_status ensure_trace_scratch();
    wf_status upload_scene();
    wf_status upload_materials_and_emitters();
    wf_status run_extend(const wf_ray* d_rays, const uint32_t* d_count, uint32_t n, wf_hit* d_hits,
                         uint32_t* d_visible, uint32_t* d_counters, bool any);
    wf_status read_overflows(uint32_t* out);
    wf_status reset_overflows();
    wf_status ensure_frame_buffers(size_t n_paths, size_t n_pixels);
    wf_status render_frame();
    void release_frame_buffers();
};
