// The renderer object (the g_renderer state of renderer_common.h:146-243,
// re-expressed over HIP device memory).
#pragma once

#include "bvh_device.h"
#include "stages.h"
#include "wf_internal.h"

#include <cstdarg>
#include <string>
#include <vector>

namespace wf {

constexpr uint32_t kMaxTextures = 65536;   // renderer.cpp:232 texture slotmap size
constexpr uint32_t kStackTotal = 128;      // LDS + spill entries per lane (reference: 64 TLAS + 64 BLAS)
constexpr int kNumStageEvents = 48;


struct Mesh {
    bool alive = false;
    uint32_t version = 0;
    uint32_t tri_count = 0, node_count = 0, max_leaf = 0;
    uint32_t root_ref = 0;     // the BLAS root's traversal reference (node 0's left_first word)
    float blas_min[3] = {0, 0, 0}, blas_max[3] = {0, 0, 0};
    std::vector<float> pos9;   // host copy of positions (emitter sampling)
    DevBuf nodes, tris, shade;
    DevBuf quads;              // 4-wide records, built on first use of that layout
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

// One scene of a frame: camera, resolution and the instances submitted to
// it (renderer::begin_scene .. end_scene makes one; a batch holds many).
struct SceneSlot {
    wf_camera camera{};
    wf_view view{};
    uint32_t width = 0, height = 0;
    std::vector<InstanceRec> instances;
    uint32_t accum_count = 0;
    // assigned by upload_scene
    uint32_t tlas_root = 0, inst_base = 0, em_off = 0, em_count = 0;
    float em_area = 0.0f;
    float bounds_min[3] = {0, 0, 0}, bounds_max[3] = {0, 0, 0};
    // assigned per frame
    uint32_t path_off = 0, pix_off = 0;
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
    uint32_t blas_builder = WF_BLAS_BUILDER_AUTO;
    wf::DevBuildScratch build_scratch;   // device BLAS builder scratch, kept across meshes
    wf::DevBuf build_pos, build_idx;
    uint32_t tlas_builder = WF_BLAS_BUILDER_AUTO;
    uint32_t tlas_metric = WF_TLAS_METRIC_REFERENCE;
    wf::DevBuf tbuild_inst, tbuild_nodes, tbuild_list, tbuild_ids, tbuild_status;
    std::string err;

    std::vector<wf::Mesh> meshes;
    std::vector<wf::Texture> textures;
    bool textures_dirty = true;

    // scene / frame state
    bool frame_open = false, scene_open = false;
    std::vector<wf::SceneSlot> slots;   // slot 0 is the begin_scene scene
    uint32_t cur_slot = 0;
    uint64_t env_texture = WF_INVALID_HANDLE;
    uint32_t xorshift_seed = 0xC977E154u;   // core/random.h:7
    uint64_t frame_index = 0;
    wf_frame_stats stats{};

    // TLAS + instance records of every slot, concatenated on the device
    wf::DevBuf d_tquads;
    wf::DevBuf d_tnodes, d_tpairs, d_inst, d_shade_inst, d_textures, d_emitters, d_scenes, d_dyn;
    bool tlas_valid = false;
    std::vector<uint8_t> h_scenes;   // staging of the per-frame scene table
    uint32_t h_dyn[4] = {0, 0, 0, 0};

    // captured frame graph
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uint64_t> graph_key;

    // the last forward frame (what the adjoint replays)
    wf::FrameArgs last_args{};
    bool forward_valid = false;
    wf::DevBuf d_lrem, d_delta, d_shadow_unit, d_shadow_light, d_adj_dl, d_adj_tex, d_adj_em;

    // traversal scratch
    wf::ExtendConfig ext_cfg;
    uint32_t ext_grid = 0, ext_grid_tlds = 0, mega_grid = 0;
    uint32_t n_tpairs = 0;
    uint32_t spill_per_thread = 0;
    wf::DevBuf d_work, d_spill, d_tmp_rays, d_tmp_hits, d_tmp_aux;
    wf::DevBuf d_order, d_order_scratch;   // cost-ordered queue of caller rays and its pre-pass scratch
    wf::DevBuf d_check;   // one word: the BLAS reference encoding's range check

    // wavefront buffers (allocated on first render)
    wf::DevBuf d_rays[2], d_paths[2], d_roots[2], d_hits, d_thr, d_path_energy, d_path_scene, d_counts;
    wf::DevBuf d_shadow_rays, d_shadow_path, d_shadow_root, d_shadow_contrib, d_shadow_visible;
    wf::DevBuf d_image, d_accum, d_final;
    size_t frame_capacity = 0;
    size_t image_pixels = 0;
    bool image_valid = false;

    wf_status fail(wf_status s, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
    wf_status hip_fail(hipError_t e, const char* what);
    wf_status encode_mesh_refs(wf::Mesh& m);
    wf::Mesh* find_mesh(uint64_t h);
    wf::Texture* find_texture(uint64_t h);
    wf_status ensure_trace_scratch();
    wf_status update_tlds_grid();
    wf_status upload_scene();
    wf_status upload_materials_and_emitters();
    // an occlusion queue traced in the same launch after a closest-hit queue
    struct OcclusionQueue {
        const wf_ray* rays;
        const uint32_t* d_count;
        uint32_t* visible;
        const uint32_t* roots;
    };
    wf_status run_extend(const wf_ray* d_rays, const uint32_t* d_count, uint32_t n, wf_hit* d_hits,
                         uint32_t* d_visible, uint32_t* d_counters, bool any, const uint32_t* d_roots = nullptr,
                         const OcclusionQueue* second = nullptr, bool reset_overflow = false,
                         const uint32_t* d_order = nullptr);
    // the cost-ordered queue of n caller rays in d_order (when the traversal
    // configuration asks for it and the launch is long enough to gain from
    // it), else *out = nullptr
    wf_status cost_order(const wf_ray* d_rays, uint32_t n, const uint32_t** out);
    wf_status read_overflows(uint32_t* out);
    wf_status reset_overflows();
    wf_status ensure_frame_buffers(size_t n_paths, size_t n_pixels);
    wf_status render_range(uint32_t first, uint32_t count);
    wf_status record_frame(const wf::FrameArgs& A, uint32_t grid, bool timed, int* n_events);
    void release_frame_buffers();
    void drop_graph();
};
