// Host-only declarations: reference wire formats and the BVH builders.
#pragma once

#include "../../include/wf.h"
#include "host_math.h"

#include <cstdint>
#include <vector>

static_assert(sizeof(wf_vertex) == 48, "vertex_t is 48 B");
static_assert(sizeof(wf_ray) == 32, "RayDesc2 is 32 B");
static_assert(sizeof(wf_hit) == 20, "hit_result_t is 20 B");
static_assert(sizeof(wf_bvh_node) == 32, "bvh_node_t is 32 B");
static_assert(sizeof(wf_bvh_triangle) == 36, "bvh_triangle_t is 36 B");
static_assert(sizeof(wf_tlas_node) == 32, "tlas_node_t is 32 B");
static_assert(sizeof(wf_bvh_instance) == 92, "bvh_instance_t is 92 B");
static_assert(sizeof(wf_view) == 272, "view_t is 272 B");

namespace wf {

// The TLAS packs two 16-bit child indices per node (tlas_builder.cpp:51,
// accelstruct.hlsl:184-185): node indices must stay below 0xFFFF.
constexpr uint32_t kMaxTlasInstances = 32767;
// BLAS stack entries pack a leaf's node index / first triangle in 27 bits.
constexpr uint32_t kMaxBlasTriangles = 1u << 26;
constexpr uint32_t kAutoDeviceBuildMin = 16384;   // create_render_mesh builds larger meshes on the GPU by default
constexpr uint32_t kAutoDeviceTlasMin = 4096;     // scenes with at least this many instances build their TLAS on the GPU by default
constexpr uint32_t kMaxInstances = 65536;   // renderer_common.h:23
constexpr uint32_t kMaxBounces = 8;         // renderer.cpp:433-441, ImGui limit :1140

// ---- host builders ----------------------------------------------------------
struct BlasData {
    std::vector<wf_bvh_node> nodes;     // node_count entries, node 1 unused
    std::vector<wf_bvh_triangle> tris;  // original order
    std::vector<uint32_t> idx;          // leaf ranges -> original triangle ids
    uint32_t node_count = 0;
    uint32_t max_leaf = 0;
};
struct TlasData {
    std::vector<wf_tlas_node> nodes;
    std::vector<wf_bvh_instance> instances;
    uint32_t node_count = 0;
};
wf_status build_blas(const float* pos9, uint32_t n, const wf_build_options& opt, BlasData& out);
wf_status build_tlas(const wf_bvh_instance* inst, uint32_t m, TlasData& out, uint32_t metric = WF_TLAS_METRIC_REFERENCE);
void blas_blob(const BlasData& d, std::vector<uint8_t>& blob);
void tlas_blob(const TlasData& d, std::vector<uint8_t>& blob);
void make_bvh_instance(const float* transform, const float* bmin, const float* bmax, uint32_t bvh_index, wf_bvh_instance& out);
void make_view(const wf_camera& cam, uint32_t width, uint32_t height, wf_view& out);

} // namespace wf
