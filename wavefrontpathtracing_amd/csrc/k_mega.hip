// Megakernel path tracer: the reference's other mode (pathtracer.hlsl:36-169,
// renderer.cpp:940-1015, settings.use_wavefront_pathtracing = 0), one
// persistent launch per frame.
//
// Every lane owns one path at a time: camera ray, closest-hit traversal,
// shading (shade_vertex, the same code as the wavefront shade stage), the
// next-event shadow ray traced any-hit in the same lane, then the
// continuation ray, until the path ends; the lane writes the path's energy
// and takes the next path.  Idle lanes are refilled from a global path
// cursor once `refill_min` of the wave's lanes are idle (as in extend).  A
// path's operations and their order are those of the wavefront frame
// (emission of a vertex, then its shadow contribution, then the next vertex),
// so both modes produce the same image bit for bit; the megakernel skips the
// ray queues, the stage launches and the stage boundaries.
#include "traverse.h"
#include "shading.h"
#include "mega.h"

#pragma clang fp contract(off)

namespace wf {
namespace {

using namespace trav;
using namespace shading;

template <int BLOCK, int LDS, bool L4, class Stack>
__device__ __forceinline__ void mega_loop(const FrameArgs& A, const MegaLaunch& M, Stack& st, uint32_t lane)
{
    const TraceScene sc = M.scene;
    const uint32_t seed = A.dyn->frame_seed;
    const uint32_t n = A.n_paths;
    TraceLane s;
    uint32_t cnt_dummy[4];
    uint32_t p = ~0u, scene = 0, root = 0, d = 0;
    bool any = false;                 // the lane's current ray is the shadow ray of bounce d
    V3 thr{1.0f, 1.0f, 1.0f};
    float4 e4 = make_float4(0, 0, 0, 0);
    float4 contrib = make_float4(0, 0, 0, 0);
    bool has_next = false;
    V3 next_o{0, 0, 0}, next_d{0, 0, 0};
    V3 cur_d{0, 0, 0};                // direction of the lane's current closest-hit ray
    uint32_t n_closest = 0, n_shadow = 0;
    bool drained = false;

    auto start = [&](V3 o, V3 dir, float tmax, bool shadow) -> bool {
        wf_ray r;
        r.origin[0] = o.x; r.origin[1] = o.y; r.origin[2] = o.z; r.tmin = kRayMinT;
        r.direction[0] = dir.x; r.direction[1] = dir.y; r.direction[2] = dir.z; r.tmax = tmax;
        any = shadow;
        if (shadow) ++n_shadow; else { ++n_closest; cur_d = dir; }
        return trace_begin<false, L4>(s, sc, r, root, cnt_dummy);
    };

    for (;;) {
        const uint64_t idle = __ballot(p == ~0u);
        const uint32_t n_idle = (uint32_t)__popcll(idle);
        if (drained && n_idle == 64u) break;
        int status = kRun;
        if (!drained && n_idle > 0u && (n_idle >= M.refill_min || n_idle == 64u)) {
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)idle) - 1u;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&M.work[0], n_idle);
            base = __shfl(base, (int)leader, 64);
            if ((uint64_t)base + n_idle >= (uint64_t)n) drained = true;
            if (p != ~0u) continue;
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
            const uint32_t i = base + rank;
            if (i >= n) continue;
            // new path: the camera ray of its pixel (generate.hlsl:12-20),
            // paths taken in the wavefront frame's first-queue order
            p = path_of_position(A, i, scene);
            const GpuScene& gs = A.scenes[scene];
            root = gs.tlas_root;
            A.path_scene[p] = scene;
            d = 0;
            thr = V3{1.0f, 1.0f, 1.0f};
            e4 = make_float4(0, 0, 0, 0);
            V3 o, dir;
            camera_ray(gs, (p - gs.path_off) % gs.n_pix, o, dir);
            if (!start(o, dir, FLT_MAX, false)) status = kDone;
        } else if (p != ~0u) {
            status = trace_pass<true, false, false, L4>(s, sc, st, nullptr, M.visits, any, cnt_dummy);
        }
        // a finished ray advances its path; the path may start another ray
        // that finishes at once (its TLAS root box missed), hence the loop
        while (status != kRun) {
            const bool overflow = status == kOverflow;
            if (overflow) {
                atomicAdd(&M.work[1], 1u);
                s.h_inst = ~0u; s.h_prim = ~0u; s.h_u = 0.0f; s.h_v = 0.0f;
            }
            status = kRun;
            bool cont = false;       // the path continues with next_o/next_d
            if (!any) {
                wf_hit hit;
                hit.instance_idx = s.h_inst; hit.primitive_idx = s.h_prim;
                hit.t = s.h_inst == ~0u ? kRayMaxT : s.t;
                hit.bary[0] = s.h_u; hit.bary[1] = s.h_v;
                VertexOut o;
                shade_vertex<false>(A, p, scene, d, seed, hit, cur_d, thr, e4, make_float4(0, 0, 0, 0), o);
                has_next = o.spawn;
                next_o = o.new_o;
                next_d = o.new_d;
                if (o.shadow) {
                    contrib = o.contrib;
                    if (!start(o.sh_o, o.sh_d, o.sh_tmax, true)) status = kDone;   // visible
                    continue;
                }
                cont = has_next;
            } else {
                // connect: add the shadow contribution when nothing blocked it
                if (s.h_inst == ~0u && !overflow) {
                    e4.x = e4.x + contrib.x; e4.y = e4.y + contrib.y; e4.z = e4.z + contrib.z;
                }
                cont = has_next;
            }
            if (cont) {
                ++d;
                if (!start(next_o, next_d, FLT_MAX, false)) status = kDone;
            } else {
                A.energy[p] = e4;
                p = ~0u;
            }
        }
    }
    // ray counts for the frame statistics: one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        n_closest += __shfl_xor(n_closest, off, 64);
        n_shadow += __shfl_xor(n_shadow, off, 64);
    }
    if (lane == 0) {
        atomicAdd(&A.counts[0], n_closest);
        atomicAdd(&A.counts[kShadowCountBase], n_shadow);
    }
}

template <int BLOCK, int LDS, bool L4>
__global__ __launch_bounds__(BLOCK, 4) void k_megakernel(FrameArgs A, MegaLaunch M)
{
    constexpr int kLdsN = LDS > 0 ? (LDS + 2) * BLOCK : 1;   // LDS levels plus the floor and the scratch level
    __shared__ uint32_t s_ent[kLdsN];
    __shared__ uint16_t s_dist[kLdsN];
    const uint32_t tid = threadIdx.x;
    const uint32_t gtid = blockIdx.x * BLOCK + tid;
    if constexpr (LDS > 0) {
        LdsStack<BLOCK, LDS, true> st{s_ent + tid, s_dist + tid, M.spill, gtid, M.spill_per_thread, gridDim.x * BLOCK, M.stack_cap};
        mega_loop<BLOCK, LDS, L4>(A, M, st, tid & 63u);
    } else {
        PrivStack<128> st;
        st.cap = M.stack_cap;
        mega_loop<BLOCK, LDS, L4>(A, M, st, tid & 63u);
    }
}

template <int BLOCK, int LDS>
void* mega_ptr(bool l4) { return l4 ? reinterpret_cast<void*>(&k_megakernel<BLOCK, LDS, true>) : reinterpret_cast<void*>(&k_megakernel<BLOCK, LDS, false>); }

void* pick_mega(uint32_t block, uint32_t lds, bool l4)
{
#define WF_CASE(B, S) if (block == B && lds == S) return mega_ptr<B, S>(l4);
    WF_CASE(64, 0) WF_CASE(64, 4) WF_CASE(64, 8) WF_CASE(64, 12) WF_CASE(64, 16)
    WF_CASE(128, 0) WF_CASE(128, 4) WF_CASE(128, 8) WF_CASE(128, 12) WF_CASE(128, 16)
    WF_CASE(256, 0) WF_CASE(256, 4) WF_CASE(256, 8) WF_CASE(256, 12) WF_CASE(256, 16)
#undef WF_CASE
    return nullptr;
}

// counts[0] and counts[kShadowCountBase] collect the frame's ray counts
__global__ void k_mega_counts_clear(uint32_t* counts)
{
    if (threadIdx.x < kCountSlots) counts[threadIdx.x] = 0u;
}

} // namespace

hipError_t mega_occupancy(const ExtendConfig& cfg, int* blocks_per_cu)
{
    void* k = pick_mega(cfg.block, cfg.lds_stack, cfg.layout == 1);
    if (!k) return hipErrorInvalidValue;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k, (int)cfg.block, 0);
}

hipError_t launch_megakernel(const FrameArgs& A, const MegaLaunch& M, const ExtendConfig& cfg, hipStream_t stream)
{
    void* k = pick_mega(cfg.block, cfg.lds_stack, cfg.layout == 1);
    if (!k) return hipErrorInvalidValue;
    k_mega_counts_clear<<<1, 64, 0, stream>>>(A.counts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    FrameArgs a = A;
    MegaLaunch m = M;
    void* args[] = {&a, &m};
    return hipLaunchKernel(k, dim3(M.grid), dim3(cfg.block), args, 0, stream);
}

} // namespace wf
