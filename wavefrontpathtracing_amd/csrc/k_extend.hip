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
//   * persistent waves pull 64-ray batches from a global atomic cursor, so
//     the grid is sized to the machine and the queue length is read on the
//     device (no indirect dispatch on HIP);
//   * one combined TLAS+BLAS stack of 8-byte entries (node reference +
//     entry distance) lives in LDS, strided by workgroup size so that lanes
//     hit distinct banks, and spills to a per-thread global region past the
//     LDS depth;
//   * stack entries carry what the reference keeps in its whole-node stack
//     (children offset, or a leaf's first triangle and count), so a pop
//     issues no memory load;
//   * a popped interior node whose entry distance is >= the current ray.t is
//     dropped without fetching its children: its children's slab entries are
//     >= its own (rounded subtraction and multiplication by the same inv_dir
//     are monotone and child boxes nest inside parent boxes), so the
//     reference would test both children, find both missing, and pop again;
//     the traversal counters still count those two child fetches;
//   * triangles are pre-permuted into leaf order with edges precomputed
//     (48-byte rows), instances are 64-byte records with 3x4 matrices.
#include "wf_internal.h"

#include <cfloat>

#pragma clang fp contract(off)

namespace wf {
namespace {

constexpr float kRayMaxT = FLT_MAX;
constexpr float kIntersectEps = 1e-8f;

__device__ __forceinline__ float hmin(float a, float b) { return fminf(a, b); }   // IEEE minNum, as HLSL min
__device__ __forceinline__ float hmax(float a, float b) { return fmaxf(a, b); }

struct Ray3 { float ox, oy, oz, dx, dy, dz, ix, iy, iz; };

__device__ __forceinline__ float slab(float mnx, float mny, float mnz, float mxx, float mxy, float mxz,
                                      const Ray3& r, float t)
{
    const float tx1 = (mnx - r.ox) * r.ix;
    const float tx2 = (mxx - r.ox) * r.ix;
    float tmin = hmin(tx1, tx2);
    float tmax = hmax(tx1, tx2);
    const float ty1 = (mny - r.oy) * r.iy;
    const float ty2 = (mxy - r.oy) * r.iy;
    tmin = hmax(tmin, hmin(ty1, ty2));
    tmax = hmin(tmax, hmax(ty1, ty2));
    const float tz1 = (mnz - r.oz) * r.iz;
    const float tz2 = (mxz - r.oz) * r.iz;
    tmin = hmax(tmin, hmin(tz1, tz2));
    tmax = hmin(tmax, hmax(tz1, tz2));
    return (tmax >= tmin && tmin < t && tmax > 0.0f) ? tmin : kRayMaxT;
}

// Moller-Trumbore with back-face culling on precomputed edges.
__device__ __forceinline__ bool tri_test(const float4& a, const float4& b, const float4& c, const Ray3& r,
                                         float& t, float& bu, float& bv)
{
    const float e1x = b.x, e1y = b.y, e1z = b.z;
    const float e2x = c.x, e2y = c.y, e2z = c.z;
    const float px = r.dy * e2z - r.dz * e2y;
    const float py = r.dz * e2x - r.dx * e2z;
    const float pz = r.dx * e2y - r.dy * e2x;
    const float det = (e1x * px + e1y * py) + e1z * pz;
    if (det < kIntersectEps) return false;
    const float inv = 1.0f / det;
    const float sx = r.ox - a.x, sy = r.oy - a.y, sz = r.oz - a.z;
    const float v = ((sx * px + sy * py) + sz * pz) * inv;
    if (v < 0.0f || v > 1.0f) return false;
    const float qx = sy * e1z - sz * e1y;
    const float qy = sz * e1x - sx * e1z;
    const float qz = sx * e1y - sy * e1x;
    const float w = ((r.dx * qx + r.dy * qy) + r.dz * qz) * inv;
    if (w < 0.0f || v + w > 1.0f) return false;
    const float tt = ((e2x * qx + e2y * qy) + e2z * qz) * inv;
    if (tt < 0.0f || tt >= t) return false;
    t = tt; bu = v; bv = w;
    return true;
}

// ---- stacks -------------------------------------------------------------------
// Entry encodings
//   TLAS interior: left_right (high half = node index >= 1, <= 0xFFFE)
//   TLAS leaf:     0xFFFF0000 | instance_idx
//   BLAS interior: left_first of the node (bit 31 clear)
//   BLAS leaf:     0x80000000 | prim_count << 27 | left_first   (prim_count 1..15, left_first < 2^27)
//                  0x80000000 | node_index                      (other leaves; the node is re-read)
template <int BLOCK, int LDS>
struct LdsStack {
    uint32_t* ent;
    float* dist;
    uint2* spill;
    uint32_t spill_n;
    uint32_t cap;
    __device__ __forceinline__ bool push(uint32_t sp, uint32_t e, float d)
    {
        if (sp >= cap) return false;
        if (sp < (uint32_t)LDS) { ent[sp * BLOCK] = e; dist[sp * BLOCK] = d; return true; }
        const uint32_t k = sp - LDS;
        if (k >= spill_n) return false;
        spill[k] = make_uint2(e, __float_as_uint(d));
        return true;
    }
    __device__ __forceinline__ void get(uint32_t sp, uint32_t& e, float& d) const
    {
        if (sp < (uint32_t)LDS) { e = ent[sp * BLOCK]; d = dist[sp * BLOCK]; return; }
        const uint2 v = spill[sp - LDS];
        e = v.x; d = __uint_as_float(v.y);
    }
};

template <int N>
struct PrivStack {
    uint32_t ent[N];
    float dist[N];
    uint32_t cap;
    __device__ __forceinline__ bool push(uint32_t sp, uint32_t e, float d)
    {
        if (sp >= (uint32_t)N || sp >= cap) return false;
        ent[sp] = e; dist[sp] = d;
        return true;
    }
    __device__ __forceinline__ void get(uint32_t sp, uint32_t& e, float& d) const { e = ent[sp]; d = dist[sp]; }
};

struct HitState { uint32_t inst, prim; float t, u, v; };

enum TraceResult { kTraceDone = 0, kTraceOverflow = 1 };

__device__ __forceinline__ float4 ld4(const void* p) { return *reinterpret_cast<const float4*>(p); }

template <bool SKIP, bool COUNT, bool ANY, class Stack>
__device__ __forceinline__ int trace_ray(const TraceScene& sc, const wf_ray& ray, uint32_t root, Stack& st, HitState& h, uint32_t* cnt)
{
    h.inst = ~0u; h.prim = ~0u; h.t = kRayMaxT; h.u = 0.0f; h.v = 0.0f;
    float t = ray.tmax;
    Ray3 wr;
    wr.ox = ray.origin[0]; wr.oy = ray.origin[1]; wr.oz = ray.origin[2];
    wr.dx = ray.direction[0]; wr.dy = ray.direction[1]; wr.dz = ray.direction[2];
    wr.ix = 1.0f / wr.dx; wr.iy = 1.0f / wr.dy; wr.iz = 1.0f / wr.dz;

    const wf_tlas_node* tn = sc.tnodes;
    const float4 ra = ld4(&tn[root]);
    const float4 rb = ld4(reinterpret_cast<const char*>(&tn[root]) + 16);
    if (COUNT) cnt[0]++;
    if (slab(ra.x, ra.y, ra.z, rb.x, rb.y, rb.z, wr, t) == kRayMaxT) return kTraceDone;
    uint32_t cur_lr = __float_as_uint(ra.w);
    uint32_t cur_ii = __float_as_uint(rb.w);

    bool in_blas = false;
    uint32_t sp = 0, base_sp = 0;
    Ray3 lr;                              // object-space ray of the current instance
    const wf_bvh_node* bn = nullptr;
    const GpuTri* bt = nullptr;
    uint32_t lf = 0, pc = 0;              // current BLAS node (left_first, prim_count)
    uint32_t inst_id = 0;

    for (;;) {
        bool pop = false;
        if (!in_blas) {
            if (cur_lr == 0) {
                // TLAS leaf: enter the instance (trace_ray_bvh_instance)
                if (COUNT) cnt[2]++;
                const GpuInstance* in = sc.inst + cur_ii;
                const float4 m0 = ld4(&in->m[0]), m1 = ld4(&in->m[4]), m2 = ld4(&in->m[8]);
                bn = in->nodes;
                bt = in->tris;
                // m[c*3+r]: m0 = (c0r0 c0r1 c0r2 c1r0), m1 = (c1r1 c1r2 c2r0 c2r1), m2 = (c2r2 c3r0 c3r1 c3r2)
                lr.ox = (m0.x * wr.ox + m0.w * wr.oy) + (m1.z * wr.oz + m2.y * 1.0f);
                lr.oy = (m0.y * wr.ox + m1.x * wr.oy) + (m1.w * wr.oz + m2.z * 1.0f);
                lr.oz = (m0.z * wr.ox + m1.y * wr.oy) + (m2.x * wr.oz + m2.w * 1.0f);
                lr.dx = (m0.x * wr.dx + m0.w * wr.dy) + (m1.z * wr.dz + m2.y * 0.0f);
                lr.dy = (m0.y * wr.dx + m1.x * wr.dy) + (m1.w * wr.dz + m2.z * 0.0f);
                lr.dz = (m0.z * wr.dx + m1.y * wr.dy) + (m2.x * wr.dz + m2.w * 0.0f);
                lr.ix = 1.0f / lr.dx; lr.iy = 1.0f / lr.dy; lr.iz = 1.0f / lr.dz;
                inst_id = cur_ii;
                lf = bn[0].left_first;
                pc = bn[0].prim_count;
                if (COUNT) cnt[1]++;
                base_sp = sp;
                in_blas = true;
                continue;
            }
            const uint32_t li = cur_lr >> 16, ri = cur_lr & 0xFFFFu;
            const float4 la = ld4(&tn[li]), lb = ld4(reinterpret_cast<const char*>(&tn[li]) + 16);
            const float4 ra = ld4(&tn[ri]), rb = ld4(reinterpret_cast<const char*>(&tn[ri]) + 16);
            if (COUNT) cnt[0] += 2;
            float dl = slab(la.x, la.y, la.z, lb.x, lb.y, lb.z, wr, t);
            float dr = slab(ra.x, ra.y, ra.z, rb.x, rb.y, rb.z, wr, t);
            uint32_t l_lr = __float_as_uint(la.w), l_ii = __float_as_uint(lb.w);
            uint32_t r_lr = __float_as_uint(ra.w), r_ii = __float_as_uint(rb.w);
            if (dl > dr) {
                float tf = dl; dl = dr; dr = tf;
                uint32_t tu = l_lr; l_lr = r_lr; r_lr = tu;
                tu = l_ii; l_ii = r_ii; r_ii = tu;
            }
            if (dl == kRayMaxT) {
                pop = true;
            } else {
                cur_lr = l_lr; cur_ii = l_ii;
                if (dr != kRayMaxT) {
                    const uint32_t e = (r_lr == 0) ? (0xFFFF0000u | r_ii) : r_lr;
                    if (!st.push(sp, e, dr)) return kTraceOverflow;
                    ++sp;
                }
            }
        } else {
            if (pc > 0) {
                for (uint32_t k = lf; k < lf + pc; ++k) {
                    const GpuTri* tr = bt + k;
                    const float4 a = ld4(&tr->v0[0]), b = ld4(&tr->e1[0]), c = ld4(&tr->e2[0]);
                    if (COUNT) cnt[3]++;
                    if (tri_test(a, b, c, lr, t, h.u, h.v)) {
                        h.prim = __float_as_uint(a.w);
                        h.inst = inst_id;
                        h.t = t;
                        if (ANY) return kTraceDone;
                    }
                }
                pop = true;
            } else {
                const float4* cp = reinterpret_cast<const float4*>(bn + lf);
                const float4 a = cp[0], b = cp[1], c = cp[2], d = cp[3];
                if (COUNT) cnt[1] += 2;
                float dl = slab(a.x, a.y, a.z, b.x, b.y, b.z, lr, t);
                float dr = slab(c.x, c.y, c.z, d.x, d.y, d.z, lr, t);
                uint32_t l_lf = __float_as_uint(a.w), l_pc = __float_as_uint(b.w), l_id = lf;
                uint32_t r_lf = __float_as_uint(c.w), r_pc = __float_as_uint(d.w), r_id = lf + 1;
                if (dl > dr) {
                    float tf = dl; dl = dr; dr = tf;
                    uint32_t tu = l_lf; l_lf = r_lf; r_lf = tu;
                    tu = l_pc; l_pc = r_pc; r_pc = tu;
                    tu = l_id; l_id = r_id; r_id = tu;
                }
                if (dl == kRayMaxT) {
                    pop = true;
                } else {
                    lf = l_lf; pc = l_pc;
                    if (dr != kRayMaxT) {
                        uint32_t e;
                        if (r_pc == 0) e = r_lf;
                        else if (r_pc < 16u && r_lf < (1u << 27)) e = 0x80000000u | (r_pc << 27) | r_lf;
                        else e = 0x80000000u | r_id;
                        if (!st.push(sp, e, dr)) return kTraceOverflow;
                        ++sp;
                    }
                }
            }
        }
        if (!pop) continue;
        // unified pop: finish the BLAS when its part of the stack is empty
        for (;;) {
            if (in_blas && sp == base_sp) in_blas = false;
            if (sp == 0) return kTraceDone;
            uint32_t e; float dd;
            --sp;
            st.get(sp, e, dd);
            if (in_blas) {
                if (e & 0x80000000u) {
                    const uint32_t cnt4 = (e >> 27) & 15u;
                    if (cnt4) { pc = cnt4; lf = e & 0x07FFFFFFu; }
                    else { const uint32_t ni = e & 0x07FFFFFFu; lf = bn[ni].left_first; pc = bn[ni].prim_count; }
                    break;
                }
                if (SKIP && dd >= t) { if (COUNT) cnt[1] += 2; continue; }
                lf = e; pc = 0;
                break;
            } else {
                if ((e >> 16) == 0xFFFFu) { cur_lr = 0; cur_ii = e & 0xFFFFu; break; }
                if (SKIP && dd >= t) { if (COUNT) cnt[0] += 2; continue; }
                cur_lr = e;
                break;
            }
        }
    }
}

template <int BLOCK, int LDS, bool SKIP, bool COUNT, bool ANY>
__global__ __launch_bounds__(BLOCK) void k_extend(ExtendLaunch L)
{
    constexpr int kLdsN = LDS > 0 ? LDS * BLOCK : 1;
    __shared__ uint32_t s_ent[kLdsN];
    __shared__ float s_dist[kLdsN];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t gtid = blockIdx.x * BLOCK + tid;
    const uint32_t n = L.d_count ? *L.d_count : L.n;

    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&L.work[0], 64u);
        base = __shfl(base, 0, 64);
        if (base >= n) break;
        const uint32_t i = base + lane;
        if (i >= n) continue;
        const wf_ray ray = L.rays[i];
        const uint32_t root = L.ray_root ? L.ray_root[i] : 0u;
        HitState h;
        uint32_t cnt[4] = {0, 0, 0, 0};
        int rc;
        if constexpr (LDS > 0) {
            LdsStack<BLOCK, LDS> st{s_ent + tid, s_dist + tid, L.spill + (size_t)gtid * L.spill_per_thread, L.spill_per_thread,
                                    L.stack_cap};
            rc = trace_ray<SKIP, COUNT, ANY>(L.scene, ray, root, st, h, cnt);
        } else {
            PrivStack<128> st;
            st.cap = L.stack_cap;
            rc = trace_ray<SKIP, COUNT, ANY>(L.scene, ray, root, st, h, cnt);
        }
        if (rc == kTraceOverflow) {
            atomicAdd(&L.work[1], 1u);
            h.inst = ~0u; h.prim = ~0u; h.t = kRayMaxT; h.u = 0.0f; h.v = 0.0f;
        }
        if (ANY) {
            L.visible[i] = (h.inst == ~0u && rc == kTraceDone) ? 1u : 0u;
        } else {
            wf_hit o;
            o.instance_idx = h.inst; o.primitive_idx = h.prim; o.t = h.t; o.bary[0] = h.u; o.bary[1] = h.v;
            L.hits[i] = o;
        }
        if (COUNT) {
            uint4 c4 = make_uint4(cnt[0], cnt[1], cnt[2], cnt[3]);
            reinterpret_cast<uint4*>(L.counters)[i] = c4;
        }
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <int BLOCK, int LDS, bool SKIP, bool COUNT, bool ANY>
void* kernel_ptr() { return reinterpret_cast<void*>(&k_extend<BLOCK, LDS, SKIP, COUNT, ANY>); }

template <bool COUNT, bool ANY>
void* pick_lds(uint32_t block, uint32_t lds, bool skip)
{
#define WF_CASE(B, S) \
    if (block == B && lds == S) return skip ? kernel_ptr<B, S, true, COUNT, ANY>() : kernel_ptr<B, S, false, COUNT, ANY>();
    WF_CASE(64, 0) WF_CASE(64, 8) WF_CASE(64, 16) WF_CASE(64, 32)
    WF_CASE(128, 0) WF_CASE(128, 8) WF_CASE(128, 16) WF_CASE(128, 32)
    WF_CASE(256, 0) WF_CASE(256, 8) WF_CASE(256, 16) WF_CASE(256, 32)
#undef WF_CASE
    return nullptr;
}

void* pick_kernel(const ExtendConfig& c, bool count, bool any)
{
    if (count) return any ? nullptr : pick_lds<true, false>(c.block, c.lds_stack, c.skip_popped != 0);
    return any ? pick_lds<false, true>(c.block, c.lds_stack, c.skip_popped != 0)
               : pick_lds<false, false>(c.block, c.lds_stack, c.skip_popped != 0);
}

} // namespace

hipError_t extend_occupancy(const ExtendConfig& cfg, int* blocks_per_cu)
{
    void* k = pick_kernel(cfg, false, false);
    if (!k) return hipErrorInvalidValue;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k, (int)cfg.block, 0);
}

hipError_t launch_extend(const ExtendLaunch& L, const ExtendConfig& cfg, hipStream_t stream)
{
    void* k = pick_kernel(cfg, L.counters != nullptr, L.any_hit != 0);
    if (!k) return hipErrorInvalidValue;
    ExtendLaunch arg = L;
    void* args[] = {&arg};
    return hipLaunchKernel(k, dim3(L.grid), dim3(cfg.block), args, 0, stream);
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint32_t n, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    k_fill_u32<<<(n + 255) / 256, 256, 0, stream>>>(p, value, n);
    return hipGetLastError();
}

} // namespace wf
