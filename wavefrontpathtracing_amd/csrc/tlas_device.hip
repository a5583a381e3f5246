// Device TLAS builder: the reference's agglomerative build
// (tlas_builder.cpp:6-68 with find_best_match :86-110, including its
// "max of mins / max of maxes" match metric) in one workgroup; UNION swaps
// in the half area of the union box as the metric.
//
// The merge sequence is the reference's and stays sequential: A's best match
// B, B's best match C, a merge when C == A, otherwise the walk moves on, with
// swap-remove of the merged pair from the candidate list.  Each best-match
// search (every remaining candidate against one node) is the parallel part: an
// argmin over the workgroup that keeps the reference's first-index tie-break,
// so the nodes, their numbering and the list order are those of the host
// build.  The walk makes a few searches per merge, each ending in barriers,
// so small instance counts run in one wave (no cross-wave reduction) and
// only large ones in 1024 lanes.
// The candidate list (node ids and bounds) lives in LDS up to kTlasLdsMax
// entries and in device memory beyond.
#include "bvh_device.h"

#include <cfloat>

#pragma clang fp contract(off)

namespace wf {
namespace {

constexpr uint32_t kTlasLdsMax = 2048;
constexpr uint32_t kTlasWaveMax = 1024;   // instance counts built by one wave

__device__ __forceinline__ float tmin_(float x, float y) { return (y < x) ? y : x; }
__device__ __forceinline__ float tmax_(float x, float y) { return (x < y) ? y : x; }
__device__ __forceinline__ float half_area3(float ex, float ey, float ez) { return ex * ey + ey * ez + ez * ex; }
// a metric's order-preserving key: -0 as +0, negative patterns inverted,
// positive ones with the sign bit set (for values below FLT_MAX, not NaN)
__device__ __forceinline__ uint32_t lane_key(float v)
{
    const uint32_t u = __float_as_uint(v == 0.0f ? 0.0f : v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// status[0] = node count, status[1] = 1 if the walk found no match or did not end
template <bool LDS, bool UNION, uint32_t kTWG>
__global__ __launch_bounds__(kTWG) void k_tlas_build(const wf_bvh_instance* __restrict__ inst, uint32_t m, wf_tlas_node* nodes,
                                                     float* g_list, uint32_t* g_ids, uint32_t* status)
{
    constexpr uint32_t CAP = LDS ? kTlasLdsMax : 1;
    __shared__ float s_list[6][CAP];
    __shared__ uint32_t s_ids[CAP];
    __shared__ uint32_t s_rk[kTWG / 64];
    __shared__ uint32_t s_ri[kTWG / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    float* lst[6];
    for (int k = 0; k < 6; ++k) lst[k] = LDS ? s_list[k] : g_list + (size_t)k * m;
    uint32_t* ids = LDS ? s_ids : g_ids;

    // leaves at nodes 1..m (tlas_builder.cpp:18-31)
    for (uint32_t i = tid; i < m; i += kTWG) {
        const wf_bvh_instance& in = inst[i];
        wf_tlas_node nd;
        for (int k = 0; k < 3; ++k) {
            nd.aabb_min[k] = in.aabb_min[k];
            nd.aabb_max[k] = in.aabb_max[k];
            lst[k][i] = in.aabb_min[k];
            lst[3 + k][i] = in.aabb_max[k];
        }
        nd.left_right = 0;
        nd.instance_idx = i;
        nodes[1 + i] = nd;
        ids[i] = 1 + i;
    }
    __syncthreads();

    // find_best_match: the first candidate of least metric, FLT_MAX and NaN never chosen
    auto best_match = [&](uint32_t a, uint32_t count) -> uint32_t {
        const float amn0 = lst[0][a], amn1 = lst[1][a], amn2 = lst[2][a];
        const float amx0 = lst[3][a], amx1 = lst[4][a], amx2 = lst[5][a];
        float best = FLT_MAX;
        uint32_t bi = 0xFFFFFFFFu;
        for (uint32_t b = tid; b < count; b += kTWG) {
            if (b == a) continue;
            float area;
            if constexpr (UNION) {
                const float l0 = tmin_(amn0, lst[0][b]), l1 = tmin_(amn1, lst[1][b]), l2 = tmin_(amn2, lst[2][b]);
                const float h0 = tmax_(amx0, lst[3][b]), h1 = tmax_(amx1, lst[4][b]), h2 = tmax_(amx2, lst[5][b]);
                area = half_area3(h0 - l0, h1 - l1, h2 - l2);
            } else {
                const float h0 = tmax_(amn0, lst[0][b]), h1 = tmax_(amn1, lst[1][b]), h2 = tmax_(amn2, lst[2][b]);
                const float l0 = tmax_(amx0, lst[3][b]), l1 = tmax_(amx1, lst[4][b]), l2 = tmax_(amx2, lst[5][b]);
                area = half_area3(h0 - l0, h1 - l1, h2 - l2);
            }
            if (area < best) { best = area; bi = b; }
        }
        // The wave's argmin in two hardware wave reductions: the least key,
        // then the least index holding it.  A key orders like the float
        // (sign-flipped bit pattern, -0 taken as +0 so that equal metrics tie
        // and the first index wins); a lane without a candidate (only FLT_MAX
        // or NaN seen) has the largest key.
        const uint32_t mine = bi != 0xFFFFFFFFu ? lane_key(best) : 0xFFFFFFFFu;
        uint32_t key = __ockl_wfred_min_u32(mine);
        bi = __ockl_wfred_min_u32(mine == key ? bi : 0xFFFFFFFFu);
        if constexpr (kTWG == 64) return bi;
        if (lane == 0) { s_rk[wave] = key; s_ri[wave] = bi; }
        __syncthreads();
        key = s_rk[0];
        bi = s_ri[0];
        for (uint32_t w = 1; w < kTWG / 64; ++w) {
            const uint32_t ok = s_rk[w], oi = s_ri[w];
            if (ok < key || (ok == key && oi < bi)) { key = ok; bi = oi; }
        }
        __syncthreads();
        return bi;
    };

    // the walk (tlas_builder.cpp:33-66); every lane holds the same state
    uint32_t count = m, at = m + 1, A = 0;
    uint32_t B = best_match(A, count);
    uint64_t guard = 0;
    const uint64_t guard_max = 4ull * m * m + 16;
    uint32_t err = 0;
    while (count > 1) {
        if (B == 0xFFFFFFFFu || ++guard > guard_max) { err = 1; break; }
        const uint32_t C = best_match(B, count);
        if (C != A) { A = B; B = C; continue; }
        if (tid == 0) {
            const uint32_t ia = ids[A], ib = ids[B];
            wf_tlas_node nn;
            nn.left_right = ia + (ib << 16);
            nn.instance_idx = 0;
            for (int k = 0; k < 3; ++k) {
                nn.aabb_min[k] = tmin_(lst[k][A], lst[k][B]);
                nn.aabb_max[k] = tmax_(lst[3 + k][A], lst[3 + k][B]);
            }
            nodes[at] = nn;
            for (int k = 0; k < 3; ++k) { lst[k][A] = nn.aabb_min[k]; lst[3 + k][A] = nn.aabb_max[k]; }
            ids[A] = at;
            for (int k = 0; k < 6; ++k) lst[k][B] = lst[k][count - 1];
            ids[B] = ids[count - 1];
        }
        __syncthreads();
        ++at;
        if (UNION && A == count - 1) A = B;   // A was the last entry and moved into B's slot
        --count;
        B = best_match(A, count);
    }
    if (tid == 0) {
        if (!err) nodes[0] = nodes[ids[A]];
        status[0] = at;
        status[1] = err;
    }
}

} // namespace

wf_status device_build_tlas(const wf_bvh_instance* d_inst, uint32_t m, wf_tlas_node* d_nodes, DevBuf& list, DevBuf& ids,
                            DevBuf& status, hipStream_t st, uint32_t& node_count, std::string& err, uint32_t metric)
{
    if (!d_inst || !d_nodes || m == 0) { err = "null pointer or no instances"; return WF_ERROR_INVALID_ARGUMENT; }
    if (metric > WF_TLAS_METRIC_UNION) { err = "unknown TLAS metric"; return WF_ERROR_INVALID_ARGUMENT; }
    const bool uni = metric == WF_TLAS_METRIC_UNION;
    if (m > kMaxTlasInstances) { err = "too many instances for 16-bit TLAS child indices"; return WF_ERROR_CAPACITY; }
    hipError_t e = status.ensure(8);
    if (e == hipSuccess && m > kTlasLdsMax) e = list.ensure(24ull * m);
    if (e == hipSuccess && m > kTlasLdsMax) e = ids.ensure(4ull * m);
    if (e != hipSuccess) { err = "TLAS builder scratch allocation failed"; return WF_ERROR_OUT_OF_MEMORY; }
    uint32_t* stp = (uint32_t*)status.p;
    if (m <= kTlasWaveMax) {
        if (uni) k_tlas_build<true, true, 64><<<1, 64, 0, st>>>(d_inst, m, d_nodes, nullptr, nullptr, stp);
        else k_tlas_build<true, false, 64><<<1, 64, 0, st>>>(d_inst, m, d_nodes, nullptr, nullptr, stp);
    } else if (m <= kTlasLdsMax) {
        if (uni) k_tlas_build<true, true, 1024><<<1, 1024, 0, st>>>(d_inst, m, d_nodes, nullptr, nullptr, stp);
        else k_tlas_build<true, false, 1024><<<1, 1024, 0, st>>>(d_inst, m, d_nodes, nullptr, nullptr, stp);
    } else {
        if (uni) k_tlas_build<false, true, 1024><<<1, 1024, 0, st>>>(d_inst, m, d_nodes, (float*)list.p, (uint32_t*)ids.p, stp);
        else k_tlas_build<false, false, 1024><<<1, 1024, 0, st>>>(d_inst, m, d_nodes, (float*)list.p, (uint32_t*)ids.p, stp);
    }
    uint32_t h[2] = {0, 0};
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h, status.p, 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) { err = std::string("TLAS build: ") + hipGetErrorString(e); return WF_ERROR_DEVICE; }
    if (h[1]) { err = "TLAS build: the match walk found no partner (non-finite bounds?)"; return WF_ERROR_INVALID_ARGUMENT; }
    node_count = h[0];
    return WF_OK;
}

} // namespace wf

extern "C" WF_API wf_status wf_tlas_build_device(const wf_bvh_instance* d_instances, uint32_t m, wf_tlas_node* d_nodes,
                                                 uint64_t node_capacity, uint32_t* out_node_count, void* stream)
{
    return wf_tlas_build_device_metric(d_instances, m, WF_TLAS_METRIC_REFERENCE, d_nodes, node_capacity, out_node_count, stream);
}

extern "C" WF_API wf_status wf_tlas_build_device_metric(const wf_bvh_instance* d_instances, uint32_t m, uint32_t metric,
                                                        wf_tlas_node* d_nodes, uint64_t node_capacity, uint32_t* out_node_count,
                                                        void* stream)
{
    if (!d_instances || !d_nodes || !out_node_count) return WF_ERROR_INVALID_ARGUMENT;
    if (node_capacity < 2ull * m) return WF_ERROR_CAPACITY;
    wf::DevBuf list, ids, status;
    std::string err;
    uint32_t count = 0;
    const wf_status s = wf::device_build_tlas(d_instances, m, d_nodes, list, ids, status, (hipStream_t)stream, count, err, metric);
    list.release();
    ids.release();
    status.release();
    if (s == WF_OK) *out_node_count = count;
    return s;
}
