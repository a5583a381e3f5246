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
// An entry is a child reference plus its slab entry distance.  Encodings:
//   interior child:  ref                      (bit 31 clear; the child's pair record index)
//   leaf child:      0x80000000 | kind << 27 | ref   (kind 1..15, ref < 2^27; TLAS leaves are kind 1)
//                    0x80000000 | node index          (BLAS leaves with more triangles or a far
//                                                      first triangle; the node is re-read)
//
// In LDS an entry takes 6 bytes: the reference plus its entry distance as a
// half-precision value rounded toward -inf.  The distance only decides the
// popped-interior skip (skip when distance >= t); a rounded-down value never
// skips a node the exact value would keep, and a node it keeps anyway is
// visited, its children both miss, and the traversal pops on exactly as the
// reference does.  Spilled entries keep the full float.
__device__ __forceinline__ uint16_t half_down(float d)
{
    const _Float16 h = (_Float16)d;                        // round to nearest even
    uint16_t b = __builtin_bit_cast(uint16_t, h);
    if ((float)h > d) b = (b & 0x8000u) ? (uint16_t)(b + 1u) : (b == 0u ? (uint16_t)0x8001u : (uint16_t)(b - 1u));
    return b;
}
__device__ __forceinline__ float half_up_to_float(uint16_t b) { return (float)__builtin_bit_cast(_Float16, b); }

// Entries past the LDS depth spill to global memory, level-major: level k of
// thread g sits at spill[k * spill_stride + g], so lanes of a wave at the same
// depth touch adjacent 8-byte words.
template <int BLOCK, int LDS>
struct LdsStack {
    uint32_t* ent;
    uint16_t* dist;
    uint2* spill;            // this thread's level-0 spill slot
    uint32_t spill_n;        // spill levels
    uint32_t spill_stride;   // threads of the grid
    uint32_t cap;
    __device__ __forceinline__ bool push(uint32_t sp, uint32_t e, float d)
    {
        if (sp >= cap) return false;
        if (sp < (uint32_t)LDS) { ent[sp * BLOCK] = e; dist[sp * BLOCK] = half_down(d); return true; }
        const uint32_t k = sp - LDS;
        if (k >= spill_n) return false;
        spill[(size_t)k * spill_stride] = make_uint2(e, __float_as_uint(d));
        return true;
    }
    __device__ __forceinline__ void get(uint32_t sp, uint32_t& e, float& d) const
    {
        if (sp < (uint32_t)LDS) { e = ent[sp * BLOCK]; d = half_up_to_float(dist[sp * BLOCK]); return; }
        const uint2 v = spill[(size_t)(sp - LDS) * spill_stride];
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

__device__ __forceinline__ float4 ld4(const void* p) { return *reinterpret_cast<const float4*>(p); }

// Loads through the global address space (the pointers below all point to
// device memory; generic pointers would compile to flat loads).
#if defined(__HIP_DEVICE_COMPILE__)
typedef float __attribute__((ext_vector_type(4))) native_float4;
typedef __attribute__((address_space(1))) const native_float4 gfloat4;
typedef __attribute__((address_space(1))) const uint32_t guint;
__device__ __forceinline__ float4 gld4(const void* p)
{
    const native_float4 v = *(gfloat4*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint32_t gldu(const void* p) { return *(guint*)p; }
#else
__device__ __forceinline__ float4 gld4(const void* p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ uint32_t gldu(const void* p) { return *reinterpret_cast<const uint32_t*>(p); }
#endif

enum LaneStatus { kRun = 0, kDone = 1, kOverflow = 2 };

// Persistent traversal with per-lane refill.  Every lane owns one ray at a
// time; one pass of the loop performs, for each lane with a ray, an interior
// visit (TLAS or BLAS: both children tested, near child next, far child
// pushed), then an instance entry or a triangle leaf if that is what the lane
// now points at, then pops until its next node is known.  `ref`/`kind` name
// the lane's node: kind 0 = interior node whose child pair sits at
// nodes + ref * 32 bytes (`nodes` = the TLAS pair array, or the BLAS node
// array, whose children are adjacent), kind != 0 = leaf (a TLAS leaf's ref
// is the instance, a BLAS leaf's ref/kind its first triangle/count).
// A lane whose ray finishes writes its result and goes idle; once at least
// `refill_min` lanes of the wave are idle, the wave claims that many rays
// from the global cursor with one atomic.  The loop ends when every lane is
// idle and the queue is drained.
template <int BLOCK, int LDS, bool SKIP, bool COUNT, bool ANY, bool TLDS, class Stack>
__device__ __forceinline__ void extend_loop(const ExtendLaunch& L, Stack& st, const float4* s_tl, uint32_t lane, uint32_t n)
{
    const TraceScene sc = L.scene;
    const uint32_t refill_min = L.refill_min;
    const uint32_t visits = L.visits;
    Ray3 wr, cr;
    float t = 0.0f;
    uint32_t ref = 0, kind = 0, sp = 0, base_sp = 0;
    bool in_blas = false;
    const char* nodes = nullptr;
    const GpuTri* tris = nullptr;
    uint32_t inst_id = 0, h_inst = ~0u, h_prim = ~0u;
    float h_u = 0.0f, h_v = 0.0f;
    uint32_t cnt[4] = {0, 0, 0, 0};
    uint32_t ray_i = ~0u;
    bool drained = false;

    for (;;) {
        const uint64_t idle = __ballot(ray_i == ~0u);
        const uint32_t n_idle = (uint32_t)__popcll(idle);
        if (drained && n_idle == 64u) break;
        int status = kRun;
        if (!drained && n_idle > 0u && (n_idle >= refill_min || n_idle == 64u)) {
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)idle) - 1u;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&L.work[0], n_idle);
            base = __shfl(base, (int)leader, 64);
            if ((uint64_t)base + n_idle >= (uint64_t)n) drained = true;
            if (ray_i != ~0u) continue;
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
            const uint32_t i = base + rank;
            if (i >= n) continue;
            // start the ray: TLAS root box tested once (trace_ray_tlas)
            ray_i = i;
            const wf_ray ray = L.rays[i];
            const uint32_t root = L.ray_root ? L.ray_root[i] : 0u;
            if (COUNT) { cnt[0] = 1; cnt[1] = 0; cnt[2] = 0; cnt[3] = 0; }
            h_inst = ~0u; h_prim = ~0u; h_u = 0.0f; h_v = 0.0f;
            t = ray.tmax;
            wr.ox = ray.origin[0]; wr.oy = ray.origin[1]; wr.oz = ray.origin[2];
            wr.dx = ray.direction[0]; wr.dy = ray.direction[1]; wr.dz = ray.direction[2];
            wr.ix = 1.0f / wr.dx; wr.iy = 1.0f / wr.dy; wr.iz = 1.0f / wr.dz;
            cr = wr;
            sp = 0; base_sp = 0; in_blas = false;
            nodes = reinterpret_cast<const char*>(sc.tpairs);
            const float4 ra = gld4(&sc.tnodes[root]);
            const float4 rb = gld4(reinterpret_cast<const char*>(&sc.tnodes[root]) + 16);
            if (slab(ra.x, ra.y, ra.z, rb.x, rb.y, rb.z, wr, t) == kRayMaxT) {
                status = kDone;
            } else {
                const uint32_t lr = __float_as_uint(ra.w);
                ref = lr == 0 ? __float_as_uint(rb.w) : 2u * __float_as_uint(rb.w);   // instance, or 2 * record
                kind = lr == 0 ? 1u : 0u;
            }
        } else if (ray_i != ~0u) {
            // pop until the next node to process is known (or the stack is empty)
            auto pop_next = [&]() {
                for (;;) {
                    if (in_blas && sp == base_sp) {
                        in_blas = false;
                        cr = wr;
                        nodes = reinterpret_cast<const char*>(sc.tpairs);
                    }
                    if (sp == 0) { status = kDone; return; }
                    uint32_t e; float dd;
                    --sp;
                    st.get(sp, e, dd);
                    if (e & 0x80000000u) {
                        const uint32_t k4 = (e >> 27) & 15u;
                        if (k4) { kind = k4; ref = e & 0x07FFFFFFu; }
                        else {
                            const char* nd = nodes + (size_t)(e & 0x07FFFFFFu) * 32u;
                            ref = gldu(nd + 12); kind = gldu(nd + 28);
                        }
                        return;
                    }
                    if (SKIP && dd >= t) { if (COUNT) cnt[in_blas ? 1 : 0] += 2; continue; }
                    ref = e; kind = 0;
                    return;
                }
            };
            // up to `visits` interior nodes per pass (each with its pops), so
            // the rarer instance and leaf steps below are shared by more work
            for (uint32_t v = 0; v < visits; ++v) {
                if (status != kRun || kind != 0) break;
                // interior node: near child next, far child pushed (swap
                // only when dist_left > dist_right)
                float4 a, b, c, d;
                if (TLDS && !in_blas) {
                    const float4* q = s_tl + 2u * ref;
                    a = q[0]; b = q[1]; c = q[2]; d = q[3];
                } else {
                    const char* cp = nodes + (size_t)ref * 32u;
                    a = gld4(cp); b = gld4(cp + 16); c = gld4(cp + 32); d = gld4(cp + 48);
                }
                if (COUNT) cnt[in_blas ? 1 : 0] += 2;
                const float dl = slab(a.x, a.y, a.z, b.x, b.y, b.z, cr, t);
                const float dr = slab(c.x, c.y, c.z, d.x, d.y, d.z, cr, t);
                const bool sw = dl > dr;
                const float dn = sw ? dr : dl, df = sw ? dl : dr;
                const uint32_t n_ref = __float_as_uint(sw ? c.w : a.w), n_kind = __float_as_uint(sw ? d.w : b.w);
                const uint32_t f_ref = __float_as_uint(sw ? a.w : c.w), f_kind = __float_as_uint(sw ? b.w : d.w);
                const uint32_t f_id = ref + (sw ? 0u : 1u);
                if (dn == kRayMaxT) {
                    pop_next();
                } else {
                    ref = n_ref; kind = n_kind;
                    if (df != kRayMaxT) {
                        uint32_t e;
                        if (f_kind == 0) e = f_ref;
                        else if (f_kind < 16u && f_ref < (1u << 27)) e = 0x80000000u | (f_kind << 27) | f_ref;
                        else e = 0x80000000u | f_id;
                        if (st.push(sp, e, df)) ++sp;
                        else status = kOverflow;
                    }
                }
            }
            if (status == kRun && kind != 0 && !in_blas) {
                // TLAS leaf: enter the instance (trace_ray_bvh_instance); the
                // BLAS root is entered without a box test
                if (COUNT) { cnt[2]++; cnt[1]++; }
                const GpuInstance* in = sc.inst + ref;
                const float4 m0 = gld4(&in->m[0]), m1 = gld4(&in->m[4]), m2 = gld4(&in->m[8]);
                const float4 pp = gld4(&in->nodes);
                const wf_bvh_node* bn;
                const GpuTri* bt;
                {
                    uint64_t u0 = (uint64_t)__float_as_uint(pp.x) | ((uint64_t)__float_as_uint(pp.y) << 32);
                    uint64_t u1 = (uint64_t)__float_as_uint(pp.z) | ((uint64_t)__float_as_uint(pp.w) << 32);
                    bn = reinterpret_cast<const wf_bvh_node*>(u0);
                    bt = reinterpret_cast<const GpuTri*>(u1);
                }
                nodes = reinterpret_cast<const char*>(bn);
                tris = bt;
                // m[c*3+r]: m0 = (c0r0 c0r1 c0r2 c1r0), m1 = (c1r1 c1r2 c2r0 c2r1), m2 = (c2r2 c3r0 c3r1 c3r2)
                cr.ox = (m0.x * wr.ox + m0.w * wr.oy) + (m1.z * wr.oz + m2.y * 1.0f);
                cr.oy = (m0.y * wr.ox + m1.x * wr.oy) + (m1.w * wr.oz + m2.z * 1.0f);
                cr.oz = (m0.z * wr.ox + m1.y * wr.oy) + (m2.x * wr.oz + m2.w * 1.0f);
                cr.dx = (m0.x * wr.dx + m0.w * wr.dy) + (m1.z * wr.dz + m2.y * 0.0f);
                cr.dy = (m0.y * wr.dx + m1.x * wr.dy) + (m1.w * wr.dz + m2.z * 0.0f);
                cr.dz = (m0.z * wr.dx + m1.y * wr.dy) + (m2.x * wr.dz + m2.w * 0.0f);
                cr.ix = 1.0f / cr.dx; cr.iy = 1.0f / cr.dy; cr.iz = 1.0f / cr.dz;
                inst_id = ref;
                ref = gldu(reinterpret_cast<const char*>(bn) + 12);    // root left_first
                kind = gldu(reinterpret_cast<const char*>(bn) + 28);   // root prim_count
                base_sp = sp;
                in_blas = true;
            }
            if (status == kRun && kind != 0 && in_blas) {
                // BLAS leaf: triangles in index order; a hit replaces the
                // closest only if t < ray.t
                for (uint32_t k = ref; k < ref + kind; ++k) {
                    const GpuTri* tr = tris + k;
                    const float4 a = gld4(&tr->v0[0]), b = gld4(&tr->e1[0]), c = gld4(&tr->e2[0]);
                    if (COUNT) cnt[3]++;
                    if (tri_test(a, b, c, cr, t, h_u, h_v)) {
                        h_prim = __float_as_uint(a.w);
                        h_inst = inst_id;
                        if (ANY) { status = kDone; break; }
                    }
                }
                if (status == kRun) pop_next();
            }
        }
        if (status != kRun) {
            if (status == kOverflow) {
                atomicAdd(&L.work[1], 1u);
                h_inst = ~0u; h_prim = ~0u; h_u = 0.0f; h_v = 0.0f;
            }
            if (ANY) {
                L.visible[ray_i] = (h_inst == ~0u && status != kOverflow) ? 1u : 0u;
            } else {
                wf_hit o;
                o.instance_idx = h_inst; o.primitive_idx = h_prim;
                o.t = h_inst == ~0u ? kRayMaxT : t;
                o.bary[0] = h_u; o.bary[1] = h_v;
                L.hits[ray_i] = o;
            }
            if (COUNT) reinterpret_cast<uint4*>(L.counters)[ray_i] = make_uint4(cnt[0], cnt[1], cnt[2], cnt[3]);
            ray_i = ~0u;
        }
    }
}

extern __shared__ float4 s_tlas_dyn[];

// TLDS: the workgroup first copies the TLAS records (at most
// kTlasLdsPairs, in dynamic LDS sized by the launch) and reads TLAS nodes
// from there.
template <int BLOCK, int LDS, bool SKIP, bool COUNT, bool ANY, bool TLDS>
__global__ __launch_bounds__(BLOCK, 8) void k_extend(ExtendLaunch L)
{
    constexpr int kLdsN = LDS > 0 ? LDS * BLOCK : 1;
    __shared__ uint32_t s_ent[kLdsN];
    __shared__ uint16_t s_dist[kLdsN];
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
        LdsStack<BLOCK, LDS> st{s_ent + tid, s_dist + tid, L.spill + gtid, L.spill_per_thread, gridDim.x * BLOCK,
                                L.stack_cap};
        extend_loop<BLOCK, LDS, SKIP, COUNT, ANY, TLDS>(L, st, s_tl, tid & 63u, n);
    } else {
        PrivStack<128> st;
        st.cap = L.stack_cap;
        extend_loop<BLOCK, LDS, SKIP, COUNT, ANY, TLDS>(L, st, s_tl, tid & 63u, n);
    }
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

void* pick_kernel(const ExtendConfig& c, bool count, bool any, bool tlds)
{
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
    void* k = pick_kernel(cfg, L.counters != nullptr, L.any_hit != 0, tlds);
    if (!k) return hipErrorInvalidValue;
    ExtendLaunch arg = L;
    void* args[] = {&arg};
    const size_t dyn = tlds ? (size_t)L.scene.n_tpairs * 64u : 0u;
    return hipLaunchKernel(k, dim3(tlds ? L.grid_tlds : L.grid), dim3(cfg.block), args, dyn, stream);
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint32_t n, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    k_fill_u32<<<(n + 255) / 256, 256, 0, stream>>>(p, value, n);
    return hipGetLastError();
}

} // namespace wf
