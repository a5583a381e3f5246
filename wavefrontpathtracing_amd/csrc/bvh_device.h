// Device BLAS builder (bvh_device.hip): the binned-SAH build of
// bvh_builder.cpp:6-168 on the GPU, with the host builder's output bytes.
#pragma once

#include <string>

#include "wf_internal.h"

namespace wf {

constexpr uint32_t kDevBuildMaxBins = 64;   // interval_count the device builder accepts (2..64)

// Device scratch of one build, grown on demand and reusable across builds.
// A fixed scratch (fixed = true) is never grown: its buffers point into
// memory the caller owns, and a build that needs more fails with
// WF_ERROR_CAPACITY; level_buf (level_cap entries) then holds the per-level
// bookkeeping the build otherwise keeps in a host vector.
struct DevBuildScratch {
    DevBuf rc[2], rl[2], rh[2];     // per-triangle records in position order (ping-pong)
    DevBuf tfr, tbl;                // partition tables, node-relative positions
    DevBuf nodes, lists, small, blocks, info, bins, chunks, cnt, off, slot_ch, ctr;
    bool fixed = false;
    uint32_t (*level_buf)[2] = nullptr;
    size_t level_cap = 0;
    void release();
};

// Bytes of every scratch buffer of a build of n triangles at B intervals, in
// the order rc, rl, rh (each twice), tfr, tbl, nodes, lists, small, blocks,
// info, bins, chunks, cnt, off, slot_ch, ctr.
constexpr int kDevBuildBuffers = 19;
void device_build_sizes(uint32_t n, uint32_t B, size_t out[kDevBuildBuffers]);

struct DevBuildResult {
    uint32_t node_count = 0;
    uint32_t max_leaf = 0;
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
    uint32_t levels = 0;            // tree levels built a level at a time
    uint32_t subtrees = 0;          // subtrees built one per wave
};

// Builds the BLAS of n triangles (9 floats each, device memory) into
// d_nodes (capacity >= 2n nodes) and d_idx (n entries).  Synchronises the
// stream a few times per tree level to size the next level's launches.
wf_status device_build_blas(const float* d_pos, uint32_t n, const wf_build_options& opt, wf_bvh_node* d_nodes,
                            uint32_t* d_idx, DevBuildScratch& S, hipStream_t stream, DevBuildResult& res,
                            std::string& err);

// The TLAS of m instance records (device memory) into d_nodes (capacity >= 2m):
// the reference's agglomerative build, node for node.  list/ids hold the
// candidate list when m exceeds what one workgroup keeps in LDS.
wf_status device_build_tlas(const wf_bvh_instance* d_inst, uint32_t m, wf_tlas_node* d_nodes, DevBuf& list, DevBuf& ids,
                            DevBuf& status, hipStream_t stream, uint32_t& node_count, std::string& err,
                            uint32_t metric = WF_TLAS_METRIC_REFERENCE);

// The 4-wide records (kQuadBytes each) of a BLAS node array, node_count / 2 of them.
// Leaves' left_first words become traversal references, in place (the
// renderer's own node copies only); *d_bad is set when a leaf's triangle range
// is out of bounds.
hipError_t launch_encode_blas_refs(wf_bvh_node* d_nodes, uint32_t node_count, GpuTri* d_tris, uint32_t tri_count,
                                   uint32_t* d_bad, hipStream_t stream);
hipError_t launch_make_quads(const wf_bvh_node* d_nodes, uint32_t node_count, void* d_quads, hipStream_t stream);

// GpuTri records in leaf order from the positions and the index array.
hipError_t launch_tris_from_index(const float* d_pos, const uint32_t* d_idx, uint32_t n, GpuTri* d_tris, hipStream_t stream);

} // namespace wf
