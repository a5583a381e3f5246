// Traversal primitives shared by the extend kernel and the megakernel:
// the slab and ray/triangle tests (intersect.hlsl:7-76), the traversal
// stacks, and one traversal pass of a lane over the TLAS -> BLAS hierarchy
// (trace_ray_tlas / trace_ray_bvh_instance / trace_ray_bvh_local,
// accelstruct.hlsl:58-218).  Device code only; FP contraction is off, so the
// results are bit-identical to that algorithm in FP32.
#pragma once

#include "wf_internal.h"

#include <cfloat>

#pragma clang fp contract(off)

namespace wf {
namespace trav {

constexpr float kRayMaxT = FLT_MAX;
constexpr float kIntersectEps = 1e-8f;

__device__ __forceinline__ float hmin(float a, float b) { return fminf(a, b); }   // IEEE minNum, as HLSL min
__device__ __forceinline__ float hmax(float a, float b) { return fmaxf(a, b); }

struct Ray3 { float ox, oy, oz, dx, dy, dz, ix, iy, iz; };

// 1/x rounded to nearest, as the IEEE division 1.0f / x gives it.  For
// |x| in [2^-126, 2^126] the hardware reciprocal estimate (within 1 ulp) is
// refined by one Newton-Raphson step with two fused multiply-adds,
// r' = r + r (1 - x r) (Markstein's reciprocal iteration); zeros,
// infinities, NaN and denormal or very large |x| take the division.
// wf_selftest_reciprocal compares the two on every float.
__device__ __forceinline__ float recip(float x)
{
    const float ax = fabsf(x);
    if (ax >= 0x1p-126f && ax <= 0x1p126f) {
        const float r = __builtin_amdgcn_rcpf(x);
        return __builtin_fmaf(__builtin_fmaf(-x, r, 1.0f), r, r);
    }
    return 1.0f / x;
}

// The three reciprocals of a direction with one branch: the fast form for
// all three, and IEEE division for all three in the lanes where any component
// lies outside [2^-126, 2^126] (both forms agree wherever the fast one is
// taken, so the results are recip()'s).
__device__ __forceinline__ void recip3(float x, float y, float z, float& rx, float& ry, float& rz)
{
    auto in_range = [](float v) { const float a = fabsf(v); return a >= 0x1p-126f && a <= 0x1p126f; };   // false for NaN
    const bool fast = in_range(x) && in_range(y) && in_range(z);
    const float fx = __builtin_amdgcn_rcpf(x), fy = __builtin_amdgcn_rcpf(y), fz = __builtin_amdgcn_rcpf(z);
    rx = __builtin_fmaf(__builtin_fmaf(-x, fx, 1.0f), fx, fx);
    ry = __builtin_fmaf(__builtin_fmaf(-y, fy, 1.0f), fy, fy);
    rz = __builtin_fmaf(__builtin_fmaf(-z, fz, 1.0f), fz, fz);
    if (!fast) {
        rx = 1.0f / x; ry = 1.0f / y; rz = 1.0f / z;
    }
}

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
    const float inv = recip(det);
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

__device__ __forceinline__ float4 ld4(const void* p) { return *reinterpret_cast<const float4*>(p); }

// Loads and stores through the global address space (the pointers below all
// point to device memory; generic pointers would compile to flat accesses).
#if defined(__HIP_DEVICE_COMPILE__)
typedef float __attribute__((ext_vector_type(4))) native_float4;
typedef uint32_t __attribute__((ext_vector_type(2))) native_uint2;
typedef __attribute__((address_space(1))) const native_float4 gfloat4;
typedef __attribute__((address_space(1))) const uint32_t guint;
typedef __attribute__((address_space(1))) const native_uint2 guint2;
typedef __attribute__((address_space(1))) uint32_t gwuint;
typedef __attribute__((address_space(1))) native_uint2 gwuint2;
__device__ __forceinline__ float4 gld4(const void* p)
{
    const native_float4 v = *(gfloat4*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint32_t gldu(const void* p) { return *(guint*)p; }
__device__ __forceinline__ uint2 gldu2(const void* p)
{
    const native_uint2 v = *(guint2*)p;
    return make_uint2(v.x, v.y);
}
__device__ __forceinline__ void gstu(void* p, uint32_t v) { *(gwuint*)p = v; }
__device__ __forceinline__ void gstu2(void* p, uint2 v)
{
    native_uint2 w;
    w.x = v.x; w.y = v.y;
    *(gwuint2*)p = w;
}
#else
__device__ __forceinline__ float4 gld4(const void* p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ uint32_t gldu(const void* p) { return *reinterpret_cast<const uint32_t*>(p); }
__device__ __forceinline__ uint2 gldu2(const void* p) { return *reinterpret_cast<const uint2*>(p); }
__device__ __forceinline__ void gstu(void* p, uint32_t v) { *reinterpret_cast<uint32_t*>(p) = v; }
__device__ __forceinline__ void gstu2(void* p, uint2 v) { *reinterpret_cast<uint2*>(p) = v; }
#endif

// In LDS the popped-interior skip also keeps each entry's distance in 16
// bits, rounded toward -inf: the distance only decides the skip (skip when
// distance >= t); a rounded-down value never skips a node the exact value
// would keep, and a node it keeps anyway is visited, its children both miss,
// and the traversal pops on exactly as the reference does.  The 16 bits are
// the upper half of the float's pattern (bfloat16): for d >= 0, dropping the
// lower half rounds toward zero, which is down; a negative distance is kept
// as -inf (it can never reach t > 0).  Spilled entries keep the full float.
__device__ __forceinline__ uint16_t dist_down(float d)
{
    return d >= 0.0f ? (uint16_t)(__float_as_uint(d) >> 16) : (uint16_t)0xFF80u;
}
__device__ __forceinline__ float dist_to_float(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// LDS stack of one thread: levels of BLOCK words (lanes of a wave hit
// distinct banks).  Level 0 is a floor and level LDS + 1 a scratch level;
// stack slot k (k < LDS) is level k + 1.  With sp entries held, level
// min(sp, LDS) holds the top entry (slot sp - 1) when sp <= LDS, and the level
// above it is where slot sp goes, or the scratch level once slots spill: a
// visit reads the top and writes its far child there unconditionally, and
// only a push advances sp.  Slots from LDS on spill to global memory,
// level-major (slot LDS + k of thread g at spill[k * spill_stride + g], so a
// wave's spills coalesce), up to `cap` entries.
template <int BLOCK, int LDS, bool DIST = false>
struct LdsStack {
    uint32_t* ent;           // this thread's level 0
    uint16_t* dist;          // entry distances (DIST), same levels
    uint2* spill;            // the grid's spill area
    uint32_t spill_id;       // this thread's column in it (its index in the grid)
    uint32_t spill_n;        // spill levels
    uint32_t spill_stride;   // threads of the grid
    uint32_t cap;
    static constexpr bool kLds = true;
    static constexpr uint32_t kLdsDepth = LDS;
    __device__ __forceinline__ uint32_t level(uint32_t sp) const { return (sp < (uint32_t)LDS ? sp : (uint32_t)LDS) * BLOCK; }
    // the top entry (valid for 0 < sp <= LDS)
    __device__ __forceinline__ uint32_t top(uint32_t sp) const { return ent[level(sp)]; }
    __device__ __forceinline__ float top_dist(uint32_t sp) const { return DIST ? dist_to_float(dist[level(sp)]) : 0.0f; }
    // the word at slot sp when sp < LDS, else the scratch level
    __device__ __forceinline__ void put(uint32_t sp, uint32_t e, float d)
    {
        ent[level(sp) + BLOCK] = e;
        if (DIST) dist[level(sp) + BLOCK] = dist_down(d);
    }
    // a push at sp >= LDS: the spill area, or false when the stack is full
    __device__ __forceinline__ bool push_spill(uint32_t sp, uint32_t e, float d)
    {
        if (sp >= cap) return false;
        const uint32_t k = sp - LDS;
        if (k >= spill_n) return false;
        uint2* q = spill + ((size_t)k * spill_stride + spill_id);
        if (DIST) gstu2(q, make_uint2(e, __float_as_uint(d)));
        else gstu(q, e);
        return true;
    }
    __device__ __forceinline__ bool push(uint32_t sp, uint32_t e, float d)
    {
        if (sp >= cap) return false;
        if (sp < (uint32_t)LDS) { put(sp, e, d); return true; }
        return push_spill(sp, e, d);
    }
    // the entry at slot sp (any depth)
    __device__ __forceinline__ void get(uint32_t sp, uint32_t& e, float& d) const
    {
        if (sp < (uint32_t)LDS) {
            e = ent[(sp + 1u) * BLOCK];
            d = DIST ? dist_to_float(dist[(sp + 1u) * BLOCK]) : 0.0f;
            return;
        }
        const uint2* q = spill + ((size_t)(sp - LDS) * spill_stride + spill_id);
        if (DIST) { const uint2 v = gldu2(q); e = v.x; d = __uint_as_float(v.y); }
        else { e = gldu(q); d = 0.0f; }
    }
};

template <int N>
struct PrivStack {
    static constexpr bool kLds = false;
    static constexpr uint32_t kLdsDepth = 0;
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

enum LaneStatus { kRun = 0, kDone = 1, kOverflow = 2 };

constexpr uint32_t kNoBlas = 0xFFFFFFFFu;

// Per-lane traversal state.  `cur` is the lane's node (encoding above):
// interior nodes are visited from `nodes` (the TLAS pair array, or the BLAS
// node array, whose children are adjacent; 4-wide records in layout 1), a
// TLAS leaf is an instance to enter, a BLAS leaf a run of triangles.  The
// world ray and the current (object-space inside a BLAS) ray share t, as in
// the reference.
struct TraceLane {
    Ray3 wr, cr;
    float t;
    uint32_t cur, sp;
    float cur_d;               // entry distance `cur` was pushed with, when it was popped (else -inf)
    uint32_t base_sp;          // stack depth at the BLAS entry; kNoBlas while in the TLAS
    __device__ __forceinline__ bool in_blas() const { return base_sp != kNoBlas; }
    const char* nodes;
    const GpuTri* tris;
    uint32_t inst_id, h_inst, h_prim;
    float h_u, h_v;
};

// Start a ray: the TLAS root box is tested once (trace_ray_tlas).  Returns
// false when the root box is missed (the ray is finished, a miss).
template <bool COUNT, bool L4 = false>
__device__ __forceinline__ bool trace_begin(TraceLane& s, const TraceScene& sc, const wf_ray& ray, uint32_t root,
                                            uint32_t* cnt)
{
    s.h_inst = ~0u; s.h_prim = ~0u; s.h_u = 0.0f; s.h_v = 0.0f;
    s.t = ray.tmax;
    s.wr.ox = ray.origin[0]; s.wr.oy = ray.origin[1]; s.wr.oz = ray.origin[2];
    s.wr.dx = ray.direction[0]; s.wr.dy = ray.direction[1]; s.wr.dz = ray.direction[2];
    recip3(s.wr.dx, s.wr.dy, s.wr.dz, s.wr.ix, s.wr.iy, s.wr.iz);
    s.cr = s.wr;
    s.sp = 0; s.base_sp = kNoBlas; s.cur_d = -INFINITY;
    s.nodes = reinterpret_cast<const char*>(L4 ? sc.tquads : sc.tpairs);
    if (COUNT) { cnt[0] = 1; cnt[1] = 0; cnt[2] = 0; cnt[3] = 0; }
    const float4 ra = gld4(&sc.tnodes[root]);
    const float4 rb = gld4(reinterpret_cast<const char*>(&sc.tnodes[root]) + 16);
    if (slab(ra.x, ra.y, ra.z, rb.x, rb.y, rb.z, s.wr, s.t) == kRayMaxT) return false;
    const uint32_t lr = __float_as_uint(ra.w);
    // the root's instance, or its record: 2 * record (pair records) / record (4-wide)
    s.cur = lr == 0 ? encode_ref(__float_as_uint(rb.w), 1u) : (L4 ? 1u : 2u) * __float_as_uint(rb.w);
    return true;
}

// One traversal pass of a lane: up to `visits` interior visits (TLAS or
// BLAS: both children tested, near child next, far child pushed; each with
// its pop), then an instance entry or a triangle leaf if that is what the
// lane now points at, then a pop.  Returns kRun, kDone (hit or miss found;
// any-hit rays stop at their first hit) or kOverflow (the stack was full).
template <bool SKIP, bool COUNT, bool TLDS, bool L4 = false, class Stack>
__device__ __forceinline__ int trace_pass(TraceLane& s, const TraceScene& sc, Stack& st, const float4* s_tl,
                                          uint32_t visits, bool any, uint32_t* cnt)
{
    static_assert(!L4 || (!COUNT && !TLDS), "the 4-wide layout has no counting or LDS-TLAS variant");
    // the LDS stack: a visit writes the far child at the top unconditionally
    // and pops with the entry it read first
    constexpr bool FAST = Stack::kLds;
    constexpr uint32_t LDSD = Stack::kLdsDepth;
    int status = kRun;
    // pushes below this depth go to LDS with no further check
    uint32_t fast_lim = 0u;
    if constexpr (FAST) fast_lim = st.cap < LDSD ? st.cap : LDSD;
    const char* tlas_nodes = reinterpret_cast<const char*>(L4 ? sc.tquads : sc.tpairs);
    // the BLAS part of the stack is used up: back to the world ray and the TLAS
    auto leave_blas = [&]() {
        s.base_sp = kNoBlas;
        s.cr = s.wr;
        s.nodes = tlas_nodes;
    };
    // pop until the next node is known (or the stack is empty)
    auto pop = [&]() {
        for (;;) {
            if (s.sp == s.base_sp) leave_blas();
            if (s.sp == 0) { status = kDone; return; }
            uint32_t e; float dd;
            --s.sp;
            st.get(s.sp, e, dd);
            if (SKIP && !(e & kLeafBit) && dd >= s.t) { if (COUNT) cnt[s.in_blas() ? 1 : 0] += 2; continue; }
            s.cur = e;
            s.cur_d = -INFINITY;
            return;
        }
    };
    auto push = [&](uint32_t e, float d) {
        if (st.push(s.sp, e, d)) ++s.sp;
        else status = kOverflow;
    };
    for (uint32_t v = 0; v < visits; ++v) {
        if (status != kRun || (s.cur & kLeafBit)) break;
        if constexpr (L4) {
            // 4-wide record: this node's children, then (same record) the
            // near child's children when it is interior
            const char* rp = s.nodes + (size_t)s.cur * kQuadBytes;
            const float4 a = gld4(rp), b = gld4(rp + 16), c = gld4(rp + 32), d = gld4(rp + 48);
            const float dl = slab(a.x, a.y, a.z, b.x, b.y, b.z, s.cr, s.t);
            const float dr = slab(c.x, c.y, c.z, d.x, d.y, d.z, s.cr, s.t);
            const bool sw = dl > dr;
            const float dn = sw ? dr : dl, df = sw ? dl : dr;
            const uint32_t en = __float_as_uint(sw ? c.w : a.w), ef = __float_as_uint(sw ? a.w : c.w);
            if (dn == kRayMaxT) { pop(); continue; }
            if (df != kRayMaxT) push(ef, df);
            if ((en & kLeafBit) || status != kRun) { s.cur = en; continue; }
            const char* np = rp + (sw ? 128u : 64u);
            const float4 e = gld4(np), f = gld4(np + 16), g = gld4(np + 32), h = gld4(np + 48);
            const float dl2 = slab(e.x, e.y, e.z, f.x, f.y, f.z, s.cr, s.t);
            const float dr2 = slab(g.x, g.y, g.z, h.x, h.y, h.z, s.cr, s.t);
            const bool sw2 = dl2 > dr2;
            const float dn2 = sw2 ? dr2 : dl2, df2 = sw2 ? dl2 : dr2;
            if (dn2 == kRayMaxT) { pop(); continue; }
            s.cur = __float_as_uint(sw2 ? g.w : e.w);
            if (df2 != kRayMaxT) push(__float_as_uint(sw2 ? e.w : g.w), df2);
            continue;
        }
        // interior node: near child next, far child pushed (swap only when
        // dist_left > dist_right)
        uint32_t pre = 0u;
        float pre_d = -INFINITY;
        if constexpr (FAST) {
            pre = st.top(s.sp);   // read before the node loads return
            if constexpr (SKIP) pre_d = st.top_dist(s.sp);
        }
        float4 a, b, c, d;
        if (TLDS && !s.in_blas()) {
            const float4* q = s_tl + 2u * s.cur;
            a = q[0]; b = q[1]; c = q[2]; d = q[3];
        } else {
            const char* cp = s.nodes + (size_t)s.cur * 32u;
            a = gld4(cp); b = gld4(cp + 16); c = gld4(cp + 32); d = gld4(cp + 48);
        }
        if (COUNT) cnt[s.in_blas() ? 1 : 0] += 2;
        const float dl = slab(a.x, a.y, a.z, b.x, b.y, b.z, s.cr, s.t);
        const float dr = slab(c.x, c.y, c.z, d.x, d.y, d.z, s.cr, s.t);
        const bool sw = dl > dr;
        // a node popped with an entry distance no closer than the current hit
        // is not opened: its children's entries are no closer, so both miss
        // (the loads are issued anyway: a branch around them costs more)
        const bool shut = SKIP && FAST && s.cur_d >= s.t;
        const float dn = shut ? kRayMaxT : (sw ? dr : dl), df = shut ? kRayMaxT : (sw ? dl : dr);
        const uint32_t en = __float_as_uint(sw ? c.w : a.w), ef = __float_as_uint(sw ? a.w : c.w);
        if constexpr (FAST) {
            // selects for the common cases; branches only for a push that
            // spills, leaving a BLAS, and a pop from an empty or spilled top
            st.put(s.sp, ef, df);
            const uint32_t sp0 = s.sp;
            const bool hit = dn != kRayMaxT, pushf = df != kRayMaxT;   // no hit: no push either
            s.cur = hit ? en : pre;
            if constexpr (SKIP) s.cur_d = hit ? -INFINITY : pre_d;
            s.sp = sp0 + (pushf ? 1u : 0u) - (hit ? 0u : 1u);
            if (!hit && sp0 == s.base_sp) leave_blas();
            if (sp0 >= fast_lim || (!hit && sp0 == 0u)) {
                // a push into the spill area or past the cap, a pop of a
                // spilled entry, or the stack is empty
                if (pushf) {
                    if (st.push(sp0, ef, df)) s.sp = sp0 + 1u;
                    else { s.sp = sp0; status = kOverflow; }
                } else if (!hit && sp0 == 0u) {
                    s.sp = 0u;
                    status = kDone;
                } else if (!hit && sp0 > LDSD) {
                    float dd;
                    st.get(sp0 - 1u, s.cur, dd);
                    if constexpr (SKIP) s.cur_d = dd;
                }
            }
        } else {
            if (dn == kRayMaxT) {
                pop();
            } else {
                s.cur = en;
                if (df != kRayMaxT) push(ef, df);
            }
        }
    }
    if (status == kRun && (s.cur & kLeafBit) && !s.in_blas()) {
        // TLAS leaf: enter the instance (trace_ray_bvh_instance); the BLAS
        // root is entered without a box test
        if (COUNT) { cnt[2]++; cnt[1]++; }
        const uint32_t id = s.cur & kRefMask;
        const GpuInstance* in = sc.inst + id;
        const float4 m0 = gld4(&in->m[0]), m1 = gld4(&in->m[4]), m2 = gld4(&in->m[8]);
        const float4 pp = gld4(&in->nodes);   // (nodes, triangles)
        const uint64_t u1 = (uint64_t)__float_as_uint(pp.z) | ((uint64_t)__float_as_uint(pp.w) << 32);
        uint64_t u0;
        uint32_t root;
        if constexpr (L4) {
            const float4 pq = gld4(&in->quads);   // (4-wide records, root2, root4)
            u0 = (uint64_t)__float_as_uint(pq.x) | ((uint64_t)__float_as_uint(pq.y) << 32);
            root = __float_as_uint(pq.w);
        } else {
            u0 = (uint64_t)__float_as_uint(pp.x) | ((uint64_t)__float_as_uint(pp.y) << 32);
            root = gldu(&in->root2);
        }
        s.nodes = reinterpret_cast<const char*>(u0);
        s.tris = reinterpret_cast<const GpuTri*>(u1);
        const Ray3& w = s.wr;
        // m[c*3+r]: m0 = (c0r0 c0r1 c0r2 c1r0), m1 = (c1r1 c1r2 c2r0 c2r1), m2 = (c2r2 c3r0 c3r1 c3r2)
        s.cr.ox = (m0.x * w.ox + m0.w * w.oy) + (m1.z * w.oz + m2.y * 1.0f);
        s.cr.oy = (m0.y * w.ox + m1.x * w.oy) + (m1.w * w.oz + m2.z * 1.0f);
        s.cr.oz = (m0.z * w.ox + m1.y * w.oy) + (m2.x * w.oz + m2.w * 1.0f);
        s.cr.dx = (m0.x * w.dx + m0.w * w.dy) + (m1.z * w.dz + m2.y * 0.0f);
        s.cr.dy = (m0.y * w.dx + m1.x * w.dy) + (m1.w * w.dz + m2.z * 0.0f);
        s.cr.dz = (m0.z * w.dx + m1.y * w.dy) + (m2.x * w.dz + m2.w * 0.0f);
        recip3(s.cr.dx, s.cr.dy, s.cr.dz, s.cr.ix, s.cr.iy, s.cr.iz);
        s.inst_id = id;
        s.cur = root;
        s.cur_d = -INFINITY;
        s.base_sp = s.sp;

    }
    if (status == kRun && (s.cur & kLeafBit) && s.in_blas()) {
        // BLAS leaf: triangles in index order; a hit replaces the closest
        // only if t < ray.t
        const uint32_t first = s.cur & kRefMask;
        uint32_t count = (s.cur >> 27) & 15u;
        if (count == 0u) count = s.tris[first].pad0;
        for (uint32_t k = first; k < first + count; ++k) {
            const GpuTri* tr = s.tris + k;
            const float4 a = gld4(&tr->v0[0]), b = gld4(&tr->e1[0]), c = gld4(&tr->e2[0]);
            if (COUNT) cnt[3]++;
            if (tri_test(a, b, c, s.cr, s.t, s.h_u, s.h_v)) {
                s.h_prim = __float_as_uint(a.w);
                s.h_inst = s.inst_id;
                if (any) { status = kDone; break; }
            }
        }
        if (status == kRun) pop();
    }
    return status;
}

} // namespace trav
} // namespace wf
