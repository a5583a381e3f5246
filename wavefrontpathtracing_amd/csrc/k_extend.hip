// Extend stage: closest-hit (and occlusion) traversal of the software TLAS ->
// BLAS hierarchy for a queue of rays.
//
// Semantics follow trace_ray_tlas / trace_ray_bvh_instance /
// trace_ray_bvh_local (accelstruct.hlsl:58-218) with intersect_ray_triangle
// and intersect_ray_aabb (intersect.hlsl:7-76), so that closest-hit ids, t and
// barycentrics are bit-identical to that algorithm in FP32 without
// contraction:
//   * TLAS root box tested once, children visited near-first (swap only when
//     dist_left > dist_right), the far child pushed unless it missed;
//   * BLAS root entered without a box test; leaf triangles in index order;
//     a triangle replaces the hit only if t < ray.t (first found keeps ties);
//   * ray.t is shared between the world and object-space rays.
//
// MI355X mechanics (not in the reference):
//   * persistent waves claim rays for their idle lanes from a queue dealt in
//     up to 64 segments with a cursor each (claims.h), so the grid is sized
//     to the machine, the queue length is read on the device (no indirect
//     dispatch on HIP) and claims are not serialised on one atomic;
//   * one combined TLAS+BLAS stack of 32-bit node references lives in LDS,
//     strided by workgroup size so that lanes hit distinct banks, and spills
//     to a per-thread global region past the LDS depth;
//   * node references are stored pre-encoded in the node records (an
//     interior child's pair, or a leaf's first triangle and count: what the
//     reference keeps in its whole-node stack), so a visit writes its far
//     child to the stack top unconditionally and a pop is a select of the
//     entry read at the start of the visit, with no memory load and no
//     divergent push or pop branch;
//   * with the opt-in popped-interior skip, entries also carry their entry
//     distance, and a popped interior node whose entry distance is >= the
//     current ray.t is dropped without fetching its children: its children's
//     slab entries are >= its own (rounded subtraction and multiplication by
//     the same inv_dir are monotone and child boxes nest inside parent boxes),
//     so the reference would test both children, find both missing, and pop
//     again; the traversal counters still count those two child fetches;
//   * triangles are pre-permuted into leaf order with edges precomputed
//     (48-byte rows), instances are 80-byte records with 3x4 matrices;
//   * an optional 4-wide node layout (L4) visits two levels per step in the
//     same order (see kQuadBytes).
#include "traverse.h"

#include <cfloat>

#pragma clang fp contract(off)

namespace wf {
namespace {

using namespace trav;

constexpr uint32_t kClaim = 256;              // most queue entries a wave claims with one atomic (BIG)
constexpr uint64_t kBigLaunch = 16ull << 20;  // queue capacity from which a mixed launch uses BIG claims

// Persistent traversal with per-lane refill.  Every lane owns one ray at a
// time and advances it one traversal pass per loop iteration (trace_pass).
// A lane whose ray finishes writes its result and goes idle; once at least
// `refill_min` lanes of the wave are idle, the wave claims that many rays
// from the queue with one atomic (ballot + mbcnt give each idle lane
// its index), and the new rays start with their root box test in the same
// iteration in which the other lanes make their pass (and then make their
// own first pass).  The loop ends when every lane is idle and the queue is
// drained.
template <int BLOCK, int LDS, bool SKIP, bool COUNT, bool ANY, bool TLDS, bool L4, bool MIX, bool BIG, class Stack>
__device__ __forceinline__ void extend_loop(const ExtendLaunch& L, Stack& st, const float4* s_tl, uint32_t lane, uint32_t n)
{
    const TraceScene sc = L.scene;
    const uint32_t refill_min = L.refill_min;
    const uint32_t visits = L.visits;
    TraceLane s;
    uint32_t cnt[4] = {0, 0, 0, 0};
    uint32_t ray_i = ~0u;
    bool drained = false;
    // MIX: a second queue of occlusion rays follows the first one
    const uint32_t n2 = MIX ? (L.d_count2 ? *L.d_count2 : L.n2) : 0u;
    const uint64_t ntot = (uint64_t)n + n2;
    bool lane_any = ANY;

    // The queue is dealt from segments with their own cursors (claims.h).
    // BIG (launches of tens of millions of entries): the wave claims up to
    // kClaim entries per atomic and keeps the unstarted ones in [lb, le); the
    // claim shrinks with the segment's entries left after the wave's previous
    // claim there, (left) / (16 x waves per segment), so ranges still held
    // when the queue runs out stay small and the tail is not lengthened.
    // Otherwise each refill claims exactly the idle lanes' entries.
    // (BIG claims are rare enough for one cursor, which keeps the band of
    // rays in flight narrowest)
    SegQueue wq;
    wq.init(L.work + kWorkClaimBase, ntot, blockIdx.x * (BLOCK / 64) + threadIdx.x / 64u, BIG ? 1u : L.claim_segs);
    uint32_t lb = 0, le = 0;   // positions of the wave's current segment
    const uint32_t share = BIG ? (16u * gridDim.x * (BLOCK / 64)) / wq.nseg + 1u : 1u;
    // the ray's result
    auto finish = [&](int status) {
        if (status == kOverflow) {
            atomicAdd(&L.work[1], 1u);
            s.h_inst = ~0u; s.h_prim = ~0u; s.h_u = 0.0f; s.h_v = 0.0f;
        }
        if (ANY) {
            L.visible[ray_i] = (s.h_inst == ~0u && status != kOverflow) ? 1u : 0u;
        } else if (MIX && lane_any) {
            L.visible2[ray_i] = (s.h_inst == ~0u && status != kOverflow) ? 1u : 0u;
        } else {
            wf_hit o;
            o.instance_idx = s.h_inst; o.primitive_idx = s.h_prim;
            o.t = s.h_inst == ~0u ? kRayMaxT : s.t;
            o.bary[0] = s.h_u; o.bary[1] = s.h_v;
            L.hits[ray_i] = o;
        }
        if (COUNT) reinterpret_cast<uint4*>(L.counters)[ray_i] = make_uint4(cnt[0], cnt[1], cnt[2], cnt[3]);
        ray_i = ~0u;
    };
    for (;;) {
        const uint64_t idle = __ballot(ray_i == ~0u);
        const uint32_t n_idle = (uint32_t)__popcll(idle);
        if (drained && n_idle == 64u) break;
        int status = kRun;
        // refill: idle lanes claim queue entries and start them; lanes that
        // hold rays make their pass in the same iteration
        if (!drained && n_idle > 0u && (n_idle >= refill_min || n_idle == 64u)) {
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)idle) - 1u;
            uint64_t q = ~0ull;
            if (lb >= le) {
                uint32_t g = n_idle;
                if constexpr (BIG) {
                    const uint32_t big = wq.left_hint() / share;
                    if (big > g) g = big < kClaim ? big : kClaim;
                }
                uint32_t first = 0;
                const uint32_t got = wq.claim(g, lane, leader, first);
                lb = first;
                le = first + got;
                if (got == 0u) drained = true;
            }
            const uint32_t take = le - lb < n_idle ? le - lb : n_idle;
            const uint32_t first = lb;
            lb += take;
            if (lb >= le && wq.empty == ~0ull) drained = true;
            if (ray_i == ~0u && rank < take) q = wq.entry(first + rank);
            if (q != ~0ull) {
                if constexpr (MIX) {
                    const bool second = q >= n;
                    const uint32_t i = second ? (uint32_t)(q - n) : (uint32_t)q;
                    ray_i = i;
                    lane_any = second;
                    const uint32_t root = second ? (L.ray_root2 ? L.ray_root2[i] : 0u) : (L.ray_root ? L.ray_root[i] : 0u);
                    if (!trace_begin<COUNT, L4>(s, sc, second ? L.rays2[i] : L.rays[i], root, cnt)) status = kDone;
                } else {
                    ray_i = L.order ? L.order[q] : (uint32_t)q;
                    const uint32_t root = L.ray_root ? L.ray_root[ray_i] : 0u;
                    if (!trace_begin<COUNT, L4>(s, sc, L.rays[ray_i], root, cnt)) status = kDone;
                }
            }
        }
        if (ray_i != ~0u && status == kRun)
            status = trace_pass<SKIP, COUNT, TLDS, L4>(s, sc, st, s_tl, visits, MIX ? lane_any : ANY, cnt);
        if (ray_i != ~0u && status != kRun) finish(status);
    }
}

extern __shared__ float4 s_tlas_dyn[];

// TLDS: the workgroup first copies the TLAS records (at most
// kTlasLdsPairs, in dynamic LDS sized by the launch) and reads TLAS nodes
// from there.
template <int BLOCK, int LDS, bool SKIP, bool COUNT, bool ANY, bool TLDS, bool L4 = false, bool MIX = false, bool BIG = false>
__global__ __launch_bounds__(BLOCK, 8) void k_extend(ExtendLaunch L)
{
    constexpr int kLdsN = LDS > 0 ? (LDS + 2) * BLOCK : 1;   // LDS levels plus the floor and the scratch level
    __shared__ uint32_t s_ent[kLdsN];
    __shared__ uint16_t s_dist[SKIP ? kLdsN : 1];   // entry distances: only the popped-interior skip reads them
    float4* s_tl = s_tlas_dyn;
    const uint32_t tid = threadIdx.x;
    const uint32_t gtid = blockIdx.x * BLOCK + tid;
    const uint32_t n = L.d_count ? *L.d_count : L.n;
    if constexpr (TLDS) {
        const uint32_t n4 = 4u * L.scene.n_tpairs;
        for (uint32_t k = tid; k < n4; k += BLOCK) s_tl[k] = gld4(L.scene.tpairs + k);
        __syncthreads();
    }
    if constexpr (LDS > 0) {
        LdsStack<BLOCK, LDS, SKIP> st{s_ent + tid, s_dist + (SKIP ? tid : 0u), L.spill, gtid, L.spill_per_thread,
                                      gridDim.x * BLOCK, L.stack_cap};
        extend_loop<BLOCK, LDS, SKIP, COUNT, ANY, TLDS, L4, MIX, BIG>(L, st, s_tl, tid & 63u, n);
    } else {
        PrivStack<128> st;
        st.cap = L.stack_cap;
        extend_loop<BLOCK, LDS, SKIP, COUNT, ANY, TLDS, L4, MIX, BIG>(L, st, s_tl, tid & 63u, n);
    }
}

__global__ void k_clear_claims(uint32_t* work, uint32_t reset_overflow)
{
    const uint32_t i = threadIdx.x;
    if (i < kClaimSegs) work[kWorkClaimBase + i * kClaimStride] = 0u;
    if (i == 0u && reset_overflow) work[1] = 0u;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <int BLOCK, int LDS, bool SKIP, bool COUNT, bool ANY, bool TLDS>
void* kernel_ptr() { return reinterpret_cast<void*>(&k_extend<BLOCK, LDS, SKIP, COUNT, ANY, TLDS>); }

template <bool COUNT, bool ANY, bool TLDS>
void* pick_lds(uint32_t block, uint32_t lds, bool skip)
{
#define WF_CASE(B, S) \
    if (block == B && lds == S) \
        return skip ? kernel_ptr<B, S, true, COUNT, ANY, TLDS>() : kernel_ptr<B, S, false, COUNT, ANY, TLDS>();
    WF_CASE(64, 0) WF_CASE(64, 4) WF_CASE(64, 8) WF_CASE(64, 12) WF_CASE(64, 16)
    WF_CASE(128, 0) WF_CASE(128, 4) WF_CASE(128, 8) WF_CASE(128, 12) WF_CASE(128, 16)
    WF_CASE(256, 0) WF_CASE(256, 4) WF_CASE(256, 8) WF_CASE(256, 12) WF_CASE(256, 16)
#undef WF_CASE
    return nullptr;
}

template <bool TLDS>
void* pick_tlds(const ExtendConfig& c, bool count, bool any)
{
    if (count) return any ? nullptr : pick_lds<true, false, TLDS>(c.block, c.lds_stack, c.skip_popped != 0);
    return any ? pick_lds<false, true, TLDS>(c.block, c.lds_stack, c.skip_popped != 0)
               : pick_lds<false, false, TLDS>(c.block, c.lds_stack, c.skip_popped != 0);
}

// 4-wide layout: popped-interior skip on, no counters, TLAS in global memory
template <bool ANY>
void* pick_quad(uint32_t block, uint32_t lds)
{
#define WF_CASE(B, S) \
    if (block == B && lds == S) return reinterpret_cast<void*>(&k_extend<B, S, true, false, ANY, false, true>);
    WF_CASE(64, 0) WF_CASE(64, 4) WF_CASE(64, 8) WF_CASE(64, 12) WF_CASE(64, 16)
    WF_CASE(128, 0) WF_CASE(128, 4) WF_CASE(128, 8) WF_CASE(128, 12) WF_CASE(128, 16)
    WF_CASE(256, 0) WF_CASE(256, 4) WF_CASE(256, 8) WF_CASE(256, 12) WF_CASE(256, 16)
#undef WF_CASE
    return nullptr;
}

// closest-hit launches with a trailing occlusion queue: no counters, TLAS in
// global memory; the 4-wide layout with the popped-interior skip only, the
// 2-wide layout with or without it and with or without BIG claims
template <bool SKIP, bool L4, bool BIG>
void* pick_mix(uint32_t block, uint32_t lds)
{
#define WF_CASE(B, S) \
    if (block == B && lds == S) return reinterpret_cast<void*>(&k_extend<B, S, SKIP, false, false, false, L4, true, BIG>);
    WF_CASE(64, 0) WF_CASE(64, 4) WF_CASE(64, 8) WF_CASE(64, 12) WF_CASE(64, 16)
    WF_CASE(128, 0) WF_CASE(128, 4) WF_CASE(128, 8) WF_CASE(128, 12) WF_CASE(128, 16)
    WF_CASE(256, 0) WF_CASE(256, 4) WF_CASE(256, 8) WF_CASE(256, 12) WF_CASE(256, 16)
#undef WF_CASE
    return nullptr;
}

void* pick_kernel(const ExtendConfig& c, bool count, bool any, bool tlds, bool mix = false, bool big = false)
{
    if (mix) {
        if (count || any || tlds) return nullptr;
        if (c.layout == 1) return c.skip_popped ? pick_mix<true, true, false>(c.block, c.lds_stack) : nullptr;
        if (c.skip_popped)
            return big ? pick_mix<true, false, true>(c.block, c.lds_stack) : pick_mix<true, false, false>(c.block, c.lds_stack);
        return big ? pick_mix<false, false, true>(c.block, c.lds_stack) : pick_mix<false, false, false>(c.block, c.lds_stack);
    }
    if (c.layout == 1) {
        if (count || tlds || !c.skip_popped) return nullptr;
        return any ? pick_quad<true>(c.block, c.lds_stack) : pick_quad<false>(c.block, c.lds_stack);
    }
    return tlds ? pick_tlds<true>(c, count, any) : pick_tlds<false>(c, count, any);
}

} // namespace

hipError_t extend_occupancy(const ExtendConfig& cfg, bool tlas_lds, size_t dyn_lds_bytes, int* blocks_per_cu)
{
    void* k = pick_kernel(cfg, false, false, tlas_lds);
    if (!k) return hipErrorInvalidValue;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k, (int)cfg.block, dyn_lds_bytes);
}

hipError_t launch_extend(const ExtendLaunch& L, const ExtendConfig& cfg, hipStream_t stream)
{
    const bool tlds = L.grid_tlds != 0;   // the host sets it when the TLAS records fit
    ExtendConfig c = cfg;
    if (L.counters) c.layout = 0;         // traversal counters come from the 2-wide kernels
    // the host knows the queues' capacities, not their lengths: capacities of
    // tens of millions mark the launches of a large batch
    const bool big = c.big == 1 || (c.big == 0 && (uint64_t)L.n + L.n2 >= kBigLaunch);
    void* k = pick_kernel(c, L.counters != nullptr, L.any_hit != 0, tlds, L.rays2 != nullptr, big);
    if (!k) return hipErrorInvalidValue;
    ExtendLaunch arg = L;
    void* args[] = {&arg};
    const size_t dyn = tlds ? (size_t)L.scene.n_tpairs * 64u : 0u;
    return hipLaunchKernel(k, dim3(tlds ? L.grid_tlds : L.grid), dim3(cfg.block), args, dyn, stream);
}

hipError_t launch_clear_claims(uint32_t* work, bool reset_overflow, hipStream_t stream)
{
    k_clear_claims<<<1, kClaimSegs, 0, stream>>>(work, reset_overflow ? 1u : 0u);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint32_t n, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    k_fill_u32<<<(n + 255) / 256, 256, 0, stream>>>(p, value, n);
    return hipGetLastError();
}

} // namespace wf
