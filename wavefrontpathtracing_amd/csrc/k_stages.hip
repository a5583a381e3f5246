// Wavefront stages around extend: clear, generate, shade (+ shadow-ray
// emission for connect), connect accumulation, resolve and post-process.
//
// Reference stage kernels: clear_buffers.hlsl:21-39, init_args.hlsl:15-22,
// generate.hlsl:12-20, shade.hlsl:42-174, post_process.hlsl:25-72.  HIP has no
// ExecuteIndirect, so every stage is a grid-stride kernel over a queue whose
// length it reads from device memory; the host never reads a count inside
// the bounce loop.  Queue appends are aggregated per wave64 (ballot +
// popcount, one atomic per wave), and the ray queues are double-buffered
// (the reference compacts in place, shade.hlsl:52 vs :167).
//
// Per-path state (throughput, energy) is FP32 and indexed by path id, so a
// path's arithmetic does not depend on where the compaction placed it; the
// random numbers are PCG32 keyed on (path, bounce).
#include "shading.h"

#include <cfloat>

#pragma clang fp contract(off)

namespace wf {
namespace {

using namespace shading;

// ---- kernels ---------------------------------------------------------------------

// clear_buffers.hlsl:21-39: throughput 1, energy 0, queue 0 holds every path
__global__ void k_clear(FrameArgs A)
{
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < A.n_paths; i += stride) {
        const uint32_t p = A.path_base + i;
        A.thr[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        A.energy[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    if (blockIdx.x == 0 && threadIdx.x < kCountSlots) A.counts[threadIdx.x] = (threadIdx.x == 0) ? A.n_paths : 0u;
}

// generate.hlsl + make_primary_ray (common.hlsl:156-169): the local path q of
// a scene covers pixel q % (W*H) (sample q / (W*H)); queue slot i holds the
// path path_of_position gives (8x8 pixel tiles within each sample's image).
__global__ void k_generate(FrameArgs A)
{
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < A.n_paths; i += stride) {
        uint32_t s;
        const uint32_t p = path_of_position(A, i, s);
        const GpuScene& sc = A.scenes[s];
        V3 o, dw;
        camera_ray(sc, (p - sc.path_off) % sc.n_pix, o, dw);
        store_ray(A.rays[0], i, o, dw, kRayMinT, FLT_MAX);
        A.path_of[0][i] = p;
        A.root_of[0][i] = sc.tlas_root;
        A.path_scene[p] = s;
    }
}

// Shade stage (shade.hlsl:42-174): one queue entry per thread, shade_vertex
// for the path, then the continuation and shadow rays appended to the next
// queues (ADJ: the path-replay adjoint of the same vertex).
//
// Each workgroup claims the next kShadeChunk entries from a cursor, so the
// queue is shaded in order (give or take the resident workgroups' chunks)
// and the appended queues keep the rays of one scene and of neighbouring
// pixels together.  A grid-stride loop lets workgroups drift apart over a
// long queue, which interleaves far apart entries (other scenes of a batch)
// in the next queue and costs the next extend launch its cache locality.
constexpr uint32_t kShadeBlock = 1024;                 // 16 waves: one append atomic per 1,024 entries
constexpr uint32_t kShadeWaves = kShadeBlock / 64u;
constexpr uint32_t kShadeChunk = 2048;

// Appends to two queues for a whole workgroup with one atomic per queue (a
// wave64 append would issue one per wave): each wave's ballot counts go to
// LDS, thread 0 reserves both ranges, and each lane's slot is the range base
// plus the counts of the waves before it plus its rank in its wave.  Every
// thread of the workgroup must call it; `par` alternates between calls so
// that a wave one call ahead writes the other half of the LDS slots.
__device__ __forceinline__ void block_append2(uint32_t* ca, bool fa, uint32_t* cb, bool fb,
                                              uint32_t (*s_cnt)[2][kShadeWaves], uint32_t (*s_rng)[2], uint32_t par,
                                              uint32_t& slot_a, uint32_t& slot_b)
{
    const uint64_t ma = __ballot(fa), mb = __ballot(fb);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) {
        s_cnt[par][0][wave] = (uint32_t)__popcll(ma);
        s_cnt[par][1][wave] = (uint32_t)__popcll(mb);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t na = 0, nb = 0;
        for (uint32_t w = 0; w < kShadeWaves; ++w) { na += s_cnt[par][0][w]; nb += s_cnt[par][1][w]; }
        s_rng[par][0] = na ? atomicAdd(ca, na) : 0u;
        s_rng[par][1] = nb ? atomicAdd(cb, nb) : 0u;
    }
    __syncthreads();
    uint32_t pa = s_rng[par][0], pb = s_rng[par][1];
    for (uint32_t w = 0; w < wave; ++w) { pa += s_cnt[par][0][w]; pb += s_cnt[par][1][w]; }
    const uint64_t below = (1ull << lane) - 1ull;
    slot_a = pa + (uint32_t)__popcll(ma & below);
    slot_b = pb + (uint32_t)__popcll(mb & below);
}

template <bool ADJ>
__global__ __launch_bounds__(kShadeBlock) void k_shade(FrameArgs A, uint32_t d)
{
    __shared__ uint32_t s_base;
    __shared__ uint32_t s_cnt[2][2][kShadeWaves];
    __shared__ uint32_t s_rng[2][2];
    uint32_t par = 0;
    const uint32_t n = A.counts[d];
    const uint32_t cur = d & 1u, nxt = cur ^ 1u;
    const uint32_t seed = A.dyn->frame_seed;
    uint32_t* cursor = A.counts + kShadeCursorBase + d;
    // workgroup-uniform control flow: the ballot-based appends below see converged lanes
    for (;;) {
        if (threadIdx.x == 0) s_base = atomicAdd(cursor, kShadeChunk);
        __syncthreads();
        const uint32_t base = s_base;
        __syncthreads();
        if (base >= n) break;
        const uint32_t stop = n - base < kShadeChunk ? n : base + kShadeChunk;
        for (uint32_t j = base; j < stop; j += kShadeBlock) {
            const uint32_t i = j + threadIdx.x;
            const bool active = i < stop;
            bool spawn = false, shadow = false;
            V3 new_o{0, 0, 0}, new_d{0, 0, 0}, sh_o{0, 0, 0}, sh_d{0, 0, 0};
            float sh_tmax = 0.0f;
            float4 contrib = make_float4(0, 0, 0, 0), unit = make_float4(0, 0, 0, 0);
            uint32_t p = 0, root = 0, light = 0;
            uint32_t em_key = 0;
            float em_g = 0.0f;
            bool em_has = false;
            if (active) {
                p = A.path_of[cur][i];
                root = A.root_of[cur][i];
                const wf_hit hit = A.hits[i];
                const float4 t4 = A.thr[p];
                float4 e4 = ADJ ? A.lrem[p] : A.energy[p];
                const float4 dlt = ADJ ? A.delta[p] : make_float4(0, 0, 0, 0);
                V3 thr{t4.x, t4.y, t4.z};
                V3 dir{0.0f, 0.0f, 0.0f};
                if (hit.instance_idx == 0xFFFFFFFFu && A.env_tex >= 0) {
                    const wf_ray& r = A.rays[cur][i];
                    dir = V3{r.direction[0], r.direction[1], r.direction[2]};
                }
                VertexOut o;
                shade_vertex<ADJ>(A, p, A.path_scene[p], d, seed, hit, dir, thr, e4, dlt, o);
                spawn = o.spawn; shadow = o.shadow;
                new_o = o.new_o; new_d = o.new_d; sh_o = o.sh_o; sh_d = o.sh_d; sh_tmax = o.sh_tmax;
                contrib = o.contrib; unit = o.unit; light = o.light;
                em_key = o.em_key; em_g = o.em_g; em_has = o.em_has;
                A.thr[p] = make_float4(thr.x, thr.y, thr.z, 0.0f);
                if (ADJ) A.lrem[p] = e4;
                else A.energy[p] = e4;
            }
            if (ADJ) wave_add_keyed(A.grad_em, em_key, em_g, em_has);
            uint32_t slot, sslot;
            block_append2(&A.counts[d + 1], spawn, &A.counts[kShadowCountBase + d], shadow, s_cnt, s_rng, par, slot, sslot);
            par ^= 1u;
            if (spawn) {
                store_ray(A.rays[nxt], slot, new_o, new_d, kRayMinT, FLT_MAX);
                A.path_of[nxt][slot] = p;
                A.root_of[nxt][slot] = root;
            }
            if (shadow) {
                store_ray(A.shadow_rays, sslot, sh_o, sh_d, kRayMinT, sh_tmax);
                A.shadow_path[sslot] = p;
                A.shadow_root[sslot] = root;
                A.shadow_contrib[sslot] = contrib;
                if (ADJ) {
                    A.shadow_unit[sslot] = unit;
                    A.shadow_light[sslot] = light;
                }
            }
        }
    }
}

// R8G8B8A8_UNORM store of one channel: NaN -> 0, clamp to [0, 1], nearest.
__device__ __forceinline__ uint32_t unorm8(float x)
{
    return (uint32_t)fminf(fmaxf(x * 255.0f + 0.5f, 0.0f), 255.0f);
}

// post_process.hlsl:25-72 after averaging the spp paths of each pixel.  The
// accumulator blends only for the path tracer's own image (:44-51); debug
// views are written as they are.  Reinhard applies to the path tracer's image
// only (:56-60) and linear -> sRGB (common.hlsl:251-259) to it and to the two
// colour views, base colour and emissive (:63-69).
__global__ void k_resolve_post(FrameArgs A)
{
    const uint32_t stride = gridDim.x * blockDim.x;
    const bool tracer = A.view_mode == WF_RENDER_VIEW_MODE_NONE;
    const bool blend = A.accumulate && tracer;
    const bool srgb = tracer || A.view_mode == WF_RENDER_VIEW_MODE_MATERIAL_BASE_COLOR ||
                      A.view_mode == WF_RENDER_VIEW_MODE_MATERIAL_EMISSIVE;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < A.n_pix; i += stride) {
        const uint32_t q = A.pix_base + i;
        const GpuScene& sc = A.scenes[scene_of_pixel(A, q)];
        const uint32_t local = q - sc.pix_off;
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        for (uint32_t s = 0; s < A.spp; ++s) {
            const float4 e = A.energy[sc.path_off + s * sc.n_pix + local];
            ax = ax + e.x; ay = ay + e.y; az = az + e.z;
        }
        const float inv = (float)A.spp;
        float4 e = make_float4(ax / inv, ay / inv, az / inv, 1.0f);
        A.image[q] = e;
        // badness detector: NaN / Inf -> magenta
        if (!(isfinite(e.x) && isfinite(e.y) && isfinite(e.z))) e = make_float4(1.0f, 0.0f, 1.0f, 1.0f);
        float4 acc;
        if (blend) {
            acc = A.accum[q];
            const float w = 1.0f / (float)sc.sample_count;
            acc.x = acc.x * (1.0f - w) + e.x * w;
            acc.y = acc.y * (1.0f - w) + e.y * w;
            acc.z = acc.z * (1.0f - w) + e.z * w;
            acc.w = acc.w * (1.0f - w) + 1.0f * w;
        } else {
            acc = e;
        }
        A.accum[q] = acc;
        float c[3] = {acc.x, acc.y, acc.z};
        for (int k = 0; k < 3; ++k) {
            if (tracer) c[k] = c[k] / (1.0f + c[k]);
            if (srgb) {
                const float x = fminf(fmaxf(c[k], 0.0f), 1.0f);
                c[k] = (x <= 0.0031308f) ? x * 12.92f : 1.055f * powf(x, 1.0f / 2.4f) - 0.055f;
            }
        }
        A.final_rgba8[q] = unorm8(c[0]) | unorm8(c[1]) << 8 | unorm8(c[2]) << 16 | unorm8(acc.w) << 24;
    }
}

// connect: add the contribution of every unoccluded shadow ray of bounce d
__global__ void k_connect_accumulate(FrameArgs A, uint32_t d)
{
    const uint32_t n = A.counts[kShadowCountBase + d];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (!A.shadow_visible[i]) continue;
        const uint32_t p = A.shadow_path[i];
        const float4 c = A.shadow_contrib[i];
        float4 e = A.energy[p];
        e.x = e.x + c.x; e.y = e.y + c.y; e.z = e.z + c.z;
        A.energy[p] = e;
    }
}

// adjoint: throughput 1, lrem = the path's forward estimate, delta = dLoss /
// d(pixel) / spp for the path's pixel
__global__ void k_adjoint_clear(FrameArgs A)
{
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < A.n_paths; i += stride) {
        const uint32_t p = A.path_base + i;
        const GpuScene& sc = A.scenes[scene_of_path(A, p)];
        const uint32_t pixel = (p - sc.path_off) % sc.n_pix;
        const float4 g = A.dl_dimage[sc.pix_off + pixel];
        const float inv = (float)A.spp;
        A.thr[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        A.lrem[p] = A.energy[p];
        A.delta[p] = make_float4(g.x / inv, g.y / inv, g.z / inv, 0.0f);
    }
    if (blockIdx.x == 0 && threadIdx.x < kCountSlots) A.counts[threadIdx.x] = (threadIdx.x == 0) ? A.n_paths : 0u;
}

// adjoint connect: collected shadow contributions leave lrem and feed the
// emission-strength gradient of their emitter
__global__ void k_adjoint_connect(FrameArgs A, uint32_t d)
{
    const uint32_t n = A.counts[kShadowCountBase + d];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t n_round = (n + 63u) & ~63u;   // whole waves for the keyed reduction
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_round; i += stride) {
        float g = 0.0f;
        uint32_t light = 0;
        bool has = false;
        if (i < n && A.shadow_visible[i]) {
            const uint32_t p = A.shadow_path[i];
            const float4 c = A.shadow_contrib[i];
            const float4 un = A.shadow_unit[i];
            const float4 dl = A.delta[p];
            float4 e = A.lrem[p];
            e.x = e.x - c.x; e.y = e.y - c.y; e.z = e.z - c.z;
            A.lrem[p] = e;
            g = (dl.x * un.x + dl.y * un.y) + dl.z * un.z;
            light = A.shadow_light[i];
            has = g != 0.0f;
        }
        wave_add_keyed(A.grad_em, light, g, has);
    }
}

} // namespace

hipError_t launch_adjoint_clear(const FrameArgs& A, uint32_t grid, hipStream_t s)
{
    k_adjoint_clear<<<grid, 256, 0, s>>>(A);
    return hipGetLastError();
}
hipError_t launch_adjoint_shade(const FrameArgs& A, uint32_t d, uint32_t grid, hipStream_t s)
{
    k_shade<true><<<grid * 256u / kShadeBlock, kShadeBlock, 0, s>>>(A, d);
    return hipGetLastError();
}
hipError_t launch_adjoint_connect(const FrameArgs& A, uint32_t d, uint32_t grid, hipStream_t s)
{
    k_adjoint_connect<<<grid, 256, 0, s>>>(A, d);
    return hipGetLastError();
}

hipError_t launch_clear(const FrameArgs& A, uint32_t grid, hipStream_t s)
{
    k_clear<<<grid, 256, 0, s>>>(A);
    return hipGetLastError();
}
hipError_t launch_generate(const FrameArgs& A, uint32_t grid, hipStream_t s)
{
    k_generate<<<grid, 256, 0, s>>>(A);
    return hipGetLastError();
}
hipError_t launch_shade(const FrameArgs& A, uint32_t d, uint32_t grid, hipStream_t s)
{
    k_shade<false><<<grid * 256u / kShadeBlock, kShadeBlock, 0, s>>>(A, d);
    return hipGetLastError();
}
hipError_t launch_connect_accumulate(const FrameArgs& A, uint32_t d, uint32_t grid, hipStream_t s)
{
    k_connect_accumulate<<<grid, 256, 0, s>>>(A, d);
    return hipGetLastError();
}
hipError_t launch_resolve_post(const FrameArgs& A, uint32_t grid, hipStream_t s)
{
    k_resolve_post<<<grid, 256, 0, s>>>(A);
    return hipGetLastError();
}

} // namespace wf
