// Internal declarations shared by the host runtime and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "claims.h"
#include "wf_host.h"

namespace wf {

// Device allocation that only grows (hipMalloc / hipFree).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n);
    void release();
};

// ---- device layouts ------------------------------------------------------------
// Triangle in leaf order (pre-permuted through the index array), with the two
// edges precomputed in FP32 (v1-v0 and v2-v0 are exactly what the
// intersection test computes first, intersect.hlsl:9-10): three 16-byte rows.
struct alignas(16) GpuTri {
    float v0[3]; uint32_t prim;    // prim = original triangle id (indices[i])
    float e1[3]; uint32_t pad0;
    float e2[3]; uint32_t pad1;
};
static_assert(sizeof(GpuTri) == 48, "GpuTri is 48 B");

// Instance record read at every TLAS leaf: rows 0..2 of world_to_local
// (column-major, m[c*3+r]), the mesh's BLAS nodes and triangles as one
// 16-byte load, and the 4-wide records (null until that layout is in use)
// with the BLAS root's reference in each layout (the root is entered without
// a box test, so no load of the root node is needed).
struct alignas(16) GpuInstance {
    float m[12];
    const wf_bvh_node* nodes;
    const GpuTri* tris;
    const char* quads;
    uint32_t root2, root4;    // the BLAS root's reference: 2-wide nodes, 4-wide records
};
static_assert(sizeof(GpuInstance) == 80, "GpuInstance is 80 B");

// Node references and stack entries of the traversal.
// A node reference (the lane's current node, and every stack entry) is one
// 32-bit word, stored pre-encoded in the .w of a child's lower half in the
// node arrays the traversal reads (BLAS nodes after kernel
// encode_blas_entries, TLAS pair records, 4-wide records):
//   interior:  ref (bit 31 clear): the child pair at nodes + ref * 32 bytes
//              (4-wide layout: the record at nodes + ref * kQuadBytes)
//   leaf:      0x80000000 | count << 27 | first, count 1..15 (a TLAS leaf:
//              count 1, first = the instance)
//              0x80000000 | first: a BLAS leaf of 16 or more triangles, whose
//              count sits in tris[first].pad0
// so a visit pushes the far child's word as it is, and a pop restores the
// node without a memory load (the reference keeps the whole node on its
// stack, accelstruct.hlsl:107-125).
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kRefMask = 0x07FFFFFFu;
constexpr uint32_t kMaxLeafRef = 1u << 27;

__host__ __device__ inline uint32_t encode_ref(uint32_t ref, uint32_t count)
{
    return count == 0u ? ref : kLeafBit | (count < 16u ? count << 27 : 0u) | ref;
}

// 4-wide records (traversal layout 1).  Record r of a hierarchy holds, for
// the child pair r of the 2-wide layout (nodes 2r and 2r+1), those two
// children followed by the children of the first and then of the second
// (zeros where a child is a leaf): six 32-byte nodes, 192 B.  An interior
// child's reference is its own record number (its left_first / 2 for a BLAS;
// its pair-record number for the TLAS).  BLAS record 0 holds the root alone.
// One visit decides the near child from the first pair and, when the near
// child is interior, its near child from that child's pair in the same
// record; every push and every box test is the 2-wide traversal's, at the
// same ray.t, so results are identical.
constexpr uint32_t kQuadBytes = 192;

// Per-triangle shading data (triangle_t without tangents): positions,
// normals and uvs of the three vertices.
struct GpuShadeTri {
    float p[3][3];
    float n[3][3];
    float uv[3][2];
};
static_assert(sizeof(GpuShadeTri) == 96, "GpuShadeTri is 96 B");

// Per-instance shading record (instance_data_t, shared.hlsl.h:87-95).
struct alignas(16) GpuShadeInstance {
    float local_to_world[16];
    float base_color[3]; int32_t base_color_tex;        // texture slot, -1 = none
    float emissive_factor[3]; int32_t emissive_tex;     // emission = emissive_factor * emissive_strength
    float emissive_strength; uint32_t pad;
    const GpuShadeTri* tris;
    int32_t normal_tex, metal_rough_tex;                 // read by the debug view modes only, -1 = none
    uint32_t pad2[2];
};
static_assert(sizeof(GpuShadeInstance) == 128, "GpuShadeInstance is 128 B");

struct GpuTexture {
    const float* texels;   // RGBA float, row-major
    uint32_t width, height;
};

// TLAS child-pair records: interior TLAS nodes are numbered densely in node
// order; record j (64 B) holds the two children of interior node j as
// {aabb_min, ref, aabb_max, kind}, with kind 0 and ref = 2 * (the child's
// record number) for an interior child (its record sits at ref * 32 bytes)
// and kind 1, ref = encode_ref(instance, 1) for a leaf; the same layout as
// two adjacent BLAS nodes {aabb_min, left_first, aabb_max, prim_count} in
// the traversal's form, so one visit routine serves both levels.  The device copy of the reference-layout
// nodes (used for the root box test) carries an interior node's record
// number in its otherwise unused instance_idx.
struct TraceScene {
    const wf_tlas_node* tnodes;
    const float4* tpairs;
    uint32_t n_tpairs;
    const GpuInstance* inst;
    const float4* tquads;     // TLAS 4-wide records (layout 1), or null
};

// most TLAS records a workgroup copies to LDS (64 B each, dynamic LDS)
constexpr uint32_t kTlasLdsPairs = 1024;

// Traversal configuration (the autotuner's axes).
struct ExtendConfig {
    // defaults: the autotuner's choice on the incoherent 256-instance workload
    uint32_t block = 256;       // workgroup size: 64, 128, 256
    uint32_t lds_stack = 16;    // LDS stack entries per lane (0 = private/scratch stack)
    uint32_t skip_popped = 1;   // skip popped interior nodes whose entry distance >= t
    uint32_t stack_limit = 0;   // cap on stack entries per lane (0 = capacity)
    uint32_t refill_min = 16;   // idle lanes of a wave that trigger a refill
    uint32_t tlas_lds = 0;      // copy the TLAS records to LDS when they fit
    uint32_t visits = 4;        // interior visits per pass of the traversal loop
    uint32_t layout = 0;        // 0: 2-wide pair records, 1: 4-wide records
    uint32_t big = 0;           // large claims in mixed launches: 0 by capacity, 1 always, 2 never
    uint32_t claim_segs = 64;   // most queue segments (cursors) of an extend launch
    uint32_t cost_order = 1;    // launches on caller rays claim the predicted-longest rays first
};

// Launch the closest-hit kernel.  d_count (device) overrides n when non-null.
struct ExtendLaunch {
    TraceScene scene;
    const wf_ray* rays;
    const uint32_t* order;     // queue entry -> ray index (a cost-ordered queue), or null for the identity
    const uint32_t* ray_root;  // per-ray TLAS root node (batched scenes), or null for root 0
    const uint32_t* d_count;
    uint32_t n;
    wf_hit* hits;              // AoS output (may be null when hit_* SoA used)
    uint32_t* visible;         // occlusion output (any-hit mode)
    uint32_t* counters;        // 4 x u32 per ray, or null
    uint32_t* work;            // [1] stack overflows; queue cursors from word kWorkClaimBase
    uint2* spill;              // global overflow stack
    uint32_t spill_per_thread;
    uint32_t stack_cap;        // entries a lane may hold (LDS + spill, or private)
    uint32_t grid;             // persistent blocks (TLAS in global memory)
    uint32_t grid_tlds;        // persistent blocks of the LDS-TLAS variant
    int any_hit;
    uint32_t refill_min;       // idle lanes of a wave that trigger a refill (1..64)
    uint32_t visits;           // interior visits per pass of the traversal loop (1..8)
    uint32_t claim_segs;       // most queue segments (a power of two, 1..64)
    // optional second queue, traced for occlusion after the first (closest-hit
    // launches only): rays2[0 .. *d_count2 or n2) with their roots, visible2 out
    const wf_ray* rays2;
    const uint32_t* ray_root2;
    const uint32_t* d_count2;
    uint32_t n2;
    uint32_t* visible2;
};
hipError_t launch_extend(const ExtendLaunch& L, const ExtendConfig& cfg, hipStream_t stream);
// zero the queue cursors of the work buffer (and its overflow word)
hipError_t launch_clear_claims(uint32_t* work, bool reset_overflow, hipStream_t stream);
hipError_t extend_occupancy(const ExtendConfig& cfg, bool tlas_lds, size_t dyn_lds_bytes, int* blocks_per_cu);
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint32_t n, hipStream_t stream);

// Cost-ordered queue of caller rays (k_order.hip): order[0, n) lists the ray
// indices with the predicted-longest traversals first, in the callers' order
// within each of kOrderBuckets cost buckets.  scratch (16-byte aligned)
// needs order_scratch_words(n) words; [bmin, bmax] is the scene's root box.
constexpr uint32_t kOrderBuckets = 16;
uint32_t order_scratch_words(uint32_t n);
hipError_t launch_cost_order(const wf_ray* rays, uint32_t n, const float bmin[3], const float bmax[3], uint32_t* scratch,
                             uint32_t* order, hipStream_t stream);

} // namespace wf
