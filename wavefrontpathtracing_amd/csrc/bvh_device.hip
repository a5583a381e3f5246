// Device BLAS builder: the reference's binned-SAH build (bvh_builder.cpp:6-168)
// on the GPU, producing the nodes and the triangle index order of the
// sequential build bit for bit (node 0 the root, node 1 unused, children
// allocated in pairs depth first with the left subtree first,
// bvh_builder.cpp:44-45,146-167).
//
// Phases:
//  * k_prep: per-triangle bounds and centroids (bvh_builder.cpp:252-264) as
//    48-byte records in position order; root bounds.
//  * one tree level at a time while nodes hold more than kSmall triangles:
//    binning (LDS bins per 2048-triangle chunk, merged with global atomics),
//    the SAH sweep of each node by one thread exactly as the reference
//    evaluates it, and the partition of every node's range.
//  * one wave per remaining subtree (<= kSmall triangles, one per lane): the
//    reference's recursion inside the wave.
//  * renumbering into the reference's depth-first pair allocation.
//
// The partition is the reference's in-place swap loop (bvh_builder.cpp:123-137:
// i runs from the front, j from the back, an element of the right side is
// swapped with a[j] and j decremented), evaluated in parallel through its
// closed form.  With n the node's size and L the size of its left side, in
// node-relative positions:
//   - a left element in [0, L) stays;
//   - the m-th right element in [0, L) (ascending) moves to n-1 if m = 1,
//     else to one below the (m-1)-th left element of [L, n) counted from the
//     back;
//   - the m-th left element of [L, n) counted from the back moves to the
//     position of the m-th right element in [0, L);
//   - a right element of [L, n) moves down by one, except the one at L, which
//     moves to one below the last left element of [L, n) counted from the
//     back (to n-1 when [L, n) holds no left element).
//
// Min and max are exact in any order, so bins and bounds take the sequential
// build's values; only the sign of a zero bound depends on which of two equal
// elements came first.  Node bounds (part of the output) therefore take the
// sign of the first zero in the node's range, in the order the reference
// scans it.  Bins, costs and split decisions compare +0 and -0 as equal, so
// they need no such care.
#include "bvh_device.h"

#include <cfloat>
#include <mutex>
#include <new>
#include <utility>
#include <vector>

#pragma clang fp contract(off)

namespace wf {
namespace {

constexpr uint32_t kSmall = 64;            // subtrees of at most this many triangles are built by one wave
constexpr uint32_t kWG = 256;
constexpr uint32_t kRows = 8;
constexpr uint32_t kChunk = kWG * kRows;   // triangles per chunk of a level pass
constexpr uint32_t kNone = 0xFFFFFFFFu;

enum : uint32_t { kOpen = 0, kLeaf = 1, kInner = 2, kSmallRoot = 3 };

// Scratch node of the build.
struct alignas(16) BNode {
    float mn[3]; uint32_t first;
    float mx[3]; uint32_t count;
    uint32_t cmn[3]; uint32_t left;        // centroid bounds as order keys
    uint32_t cmx[3]; uint32_t right;
    uint32_t kind, desc, fid, cursor;      // desc: descendants; fid: final index; cursor: final index of its first child
    uint32_t block, buf, inner, pad0;      // small roots: subtree node block, record buffer, root is interior
    uint32_t fz[6]; uint32_t pad1[2];      // first position holding a zero bound (signed-zero fix-up)
};
static_assert(sizeof(BNode) == 128, "BNode is 128 B");

// Split of one node of the current level.
struct SplitInfo {
    uint32_t node, mode, axis, split;      // mode: 0 leaf, 1 split, 2 partition then leaf
    uint32_t L, pad0;
    float lo, scale;
};

struct Counters {
    uint32_t nodes, next, small, blocks;
    uint32_t negzero, nonfinite, max_leaf, bad_index;   // bad_index: an internal index check failed
    uint32_t root_keys[12];                // vertex min, vertex max, centroid min, centroid max
};

// ---- helpers ------------------------------------------------------------------

__device__ __forceinline__ float gmin(float x, float y) { return (y < x) ? y : x; }
__device__ __forceinline__ float gmax(float x, float y) { return (x < y) ? y : x; }
__device__ __forceinline__ float half_area(float ex, float ey, float ez) { return ex * ey + ey * ez + ez * ex; }

// order-preserving key of a float (-0 orders below +0)
__device__ __forceinline__ uint32_t fkey(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float kfloat(uint32_t k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}
#define KEY_FMAX 0xFF7FFFFFu    // fkey(FLT_MAX)
#define KEY_NFMAX 0x00800000u   // fkey(-FLT_MAX)

// bin of a centroid, min(B-1, int((c - lo) * scale)) (bvh_builder.cpp:194-195);
// saturating where the reference's conversion would be undefined
__device__ __forceinline__ uint32_t bin_of(float c, float lo, float scale, uint32_t bins)
{
    const float f = (c - lo) * scale;
    if (!(f >= 0.0f)) return 0;
    if (f >= (float)bins) return bins - 1;
    const uint32_t b = (uint32_t)(int32_t)f;
    return b < bins - 1 ? b : bins - 1;
}

__device__ __forceinline__ float rec_get(const float4& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// Wave-wide reductions and inclusive scans over all 64 lanes (every call
// site runs with the whole wave active), from the ROCm device library: DPP
// row operations instead of LDS permutes.
extern "C" __device__ uint32_t __ockl_wfred_min_u32(uint32_t);
extern "C" __device__ uint32_t __ockl_wfred_max_u32(uint32_t);
extern "C" __device__ uint32_t __ockl_wfred_add_u32(uint32_t);
extern "C" __device__ uint64_t __ockl_wfred_min_u64(uint64_t);
extern "C" __device__ uint32_t __ockl_wfscan_add_u32(uint32_t, bool);
extern "C" __device__ uint32_t __ockl_wfscan_min_u32(uint32_t, bool);
extern "C" __device__ uint32_t __ockl_wfscan_max_u32(uint32_t, bool);

__device__ __forceinline__ uint32_t wave_min_u(uint32_t v) { return __ockl_wfred_min_u32(v); }
__device__ __forceinline__ uint32_t wave_max_u(uint32_t v) { return __ockl_wfred_max_u32(v); }
__device__ __forceinline__ uint32_t wave_sum_u(uint32_t v) { return __ockl_wfred_add_u32(v); }
__device__ __forceinline__ uint32_t lanes_below(uint64_t m, uint32_t lane)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}
__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Exclusive block-wide prefix of two flags in thread order; totals returned.
__device__ __forceinline__ void block_scan2(bool f0, bool f1, uint32_t& ex0, uint32_t& ex1, uint32_t& t0, uint32_t& t1,
                                            uint32_t* s_w)
{
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t b0 = __ballot(f0), b1 = __ballot(f1);
    if (lane == 0) { s_w[2 * wave] = (uint32_t)__popcll(b0); s_w[2 * wave + 1] = (uint32_t)__popcll(b1); }
    __syncthreads();
    uint32_t p0 = 0, p1 = 0;
    t0 = 0; t1 = 0;
    for (uint32_t w = 0; w < kWG / 64; ++w) {
        const uint32_t a = s_w[2 * w], b = s_w[2 * w + 1];
        if (w < wave) { p0 += a; p1 += b; }
        t0 += a; t1 += b;
    }
    __syncthreads();
    ex0 = p0 + lanes_below(b0, lane);
    ex1 = p1 + lanes_below(b1, lane);
}

// ---- per-triangle records and the root ------------------------------------------

// Grid-stride over the triangles; the root bounds are reduced per lane, per
// wave and per workgroup before one set of atomics per workgroup.
__global__ __launch_bounds__(256) void k_prep(const float* __restrict__ pos, uint32_t n, float4* rc, float4* rl, float4* rh,
                                              Counters* ctr)
{
    __shared__ uint32_t s_k[4][14];
    uint32_t keys[12];
    for (int k = 0; k < 3; ++k) {
        keys[k] = KEY_FMAX; keys[3 + k] = KEY_NFMAX; keys[6 + k] = KEY_FMAX; keys[9 + k] = KEY_NFMAX;
    }
    uint32_t bad = 0, negz = 0;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const float* p = pos + 9ull * t;
        float mn[3], mx[3], c[3];
        for (int k = 0; k < 3; ++k) {
            const float a = p[k], b = p[3 + k], d = p[6 + k];
            if (!isfinite(a) || !isfinite(b) || !isfinite(d)) bad = 1;
            // get_triangle_min_max / get_triangle_centroid (bvh_builder.cpp:252-264)
            mn[k] = gmin(gmin(a, b), d);
            mx[k] = gmax(gmax(a, b), d);
            c[k] = ((a + b) + d) * 0.3333f;
            if (__float_as_uint(mn[k]) == 0x80000000u || __float_as_uint(mx[k]) == 0x80000000u) negz = 1;
            const uint32_t kl = fkey(mn[k]), kh = fkey(mx[k]), kc = fkey(c[k]);
            keys[k] = kl < keys[k] ? kl : keys[k];
            keys[3 + k] = kh > keys[3 + k] ? kh : keys[3 + k];
            keys[6 + k] = kc < keys[6 + k] ? kc : keys[6 + k];
            keys[9 + k] = kc > keys[9 + k] ? kc : keys[9 + k];
        }
        rc[t] = make_float4(c[0], c[1], c[2], __uint_as_float(t));
        rl[t] = make_float4(mn[0], mn[1], mn[2], 0.0f);
        rh[t] = make_float4(mx[0], mx[1], mx[2], 0.0f);
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (int k = 0; k < 3; ++k) {
        keys[k] = wave_min_u(keys[k]);
        keys[3 + k] = wave_max_u(keys[3 + k]);
        keys[6 + k] = wave_min_u(keys[6 + k]);
        keys[9 + k] = wave_max_u(keys[9 + k]);
    }
    bad = wave_max_u(bad);
    negz = wave_max_u(negz);
    if (lane == 0) {
        for (int k = 0; k < 12; ++k) s_k[wave][k] = keys[k];
        s_k[wave][12] = bad;
        s_k[wave][13] = negz;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < blockDim.x / 64u; ++w)
            for (int k = 0; k < 14; ++k) {
                const uint32_t u = s_k[w][k];
                const bool take_min = (k % 6) < 3 && k < 12;
                s_k[0][k] = k >= 12 ? (s_k[0][k] | u) : (take_min ? (u < s_k[0][k] ? u : s_k[0][k]) : (u > s_k[0][k] ? u : s_k[0][k]));
            }
        for (int k = 0; k < 3; ++k) {
            atomicMin(&ctr->root_keys[k], s_k[0][k]);
            atomicMax(&ctr->root_keys[3 + k], s_k[0][3 + k]);
            atomicMin(&ctr->root_keys[6 + k], s_k[0][6 + k]);
            atomicMax(&ctr->root_keys[9 + k], s_k[0][9 + k]);
        }
        if (s_k[0][12]) atomicOr(&ctr->nonfinite, 1u);
        if (s_k[0][13]) atomicOr(&ctr->negzero, 1u);
    }
}

__global__ void k_ctr_init(Counters* ctr)
{
    const uint32_t i = threadIdx.x;
    if (i < 8) (&ctr->nodes)[i] = 0u;   // nodes .. bad_index
    if (i < 12) ctr->root_keys[i] = ((i % 6) < 3) ? KEY_FMAX : KEY_NFMAX;
}

// Node 0 of the scratch tree is the root (first 0, count n).
__global__ void k_root_init(BNode* nodes, uint32_t n, Counters* ctr, uint32_t* lists, uint32_t* small)
{
    BNode& r = nodes[0];
    for (int k = 0; k < 3; ++k) {
        r.mn[k] = kfloat(ctr->root_keys[k]);
        r.mx[k] = kfloat(ctr->root_keys[3 + k]);
        r.cmn[k] = ctr->root_keys[6 + k];
        r.cmx[k] = ctr->root_keys[9 + k];
    }
    for (int k = 0; k < 6; ++k) r.fz[k] = kNone;
    r.first = 0;
    r.count = n;
    r.left = r.right = kNone;
    r.desc = 0;
    r.fid = 0;
    r.cursor = 2;
    r.block = 0;
    r.buf = 0;
    r.inner = 0;
    ctr->nodes = 1;
    if (n <= kSmall) {
        r.kind = kSmallRoot;
        small[0] = 0;
        ctr->small = 1;
    } else {
        r.kind = kOpen;
        lists[0] = 0;
    }
}

// ---- chunk lists ------------------------------------------------------------------

__global__ void k_chunk_count(const uint32_t* list, uint32_t a, const BNode* nodes, uint2* cnt)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a) return;
    cnt[s] = make_uint2((nodes[list[s]].count + kChunk - 1) / kChunk, 0u);
}

// Exclusive prefix of m uint2 entries into off[0..m]; one workgroup.
__global__ __launch_bounds__(1024) void k_scan(const uint2* in, uint32_t m, uint2* off)
{
    __shared__ uint32_t s_w[2 * 16];
    __shared__ uint2 s_carry;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    if (tid == 0) s_carry = make_uint2(0u, 0u);
    __syncthreads();
    for (uint32_t base = 0; base < m; base += 4096u) {
        uint2 v[4];
        uint32_t sx = 0, sy = 0;
        for (int j = 0; j < 4; ++j) {
            const uint32_t i = base + tid * 4u + (uint32_t)j;
            v[j] = i < m ? in[i] : make_uint2(0u, 0u);
            sx += v[j].x;
            sy += v[j].y;
        }
        // inclusive wave scan of the thread sums
        uint32_t ix = sx, iy = sy;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t tx = __shfl_up(ix, o, 64), ty = __shfl_up(iy, o, 64);
            if (lane >= (uint32_t)o) { ix += tx; iy += ty; }
        }
        if (lane == 63) { s_w[2 * wave] = ix; s_w[2 * wave + 1] = iy; }
        __syncthreads();
        uint32_t px = s_carry.x, py = s_carry.y, tx = 0, ty = 0;
        for (uint32_t w = 0; w < 16; ++w) {
            if (w < wave) { px += s_w[2 * w]; py += s_w[2 * w + 1]; }
            tx += s_w[2 * w];
            ty += s_w[2 * w + 1];
        }
        uint32_t ex = px + ix - sx, ey = py + iy - sy;
        for (int j = 0; j < 4; ++j) {
            const uint32_t i = base + tid * 4u + (uint32_t)j;
            if (i < m) off[i] = make_uint2(ex, ey);
            ex += v[j].x;
            ey += v[j].y;
        }
        __syncthreads();
        if (tid == 0) s_carry = make_uint2(s_carry.x + tx, s_carry.y + ty);
        __syncthreads();
    }
    if (tid == 0) off[m] = s_carry;
}

__global__ void k_chunk_fill(const uint32_t* list, uint32_t a, const BNode* nodes, const uint2* off, uint4* chunks, uint2* slot_ch)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a) return;
    const BNode& nd = nodes[list[s]];
    const uint32_t c0 = off[s].x, nch = off[s + 1].x - c0;
    slot_ch[s] = make_uint2(c0, nch);
    const uint32_t end = nd.first + nd.count;
    for (uint32_t k = 0; k < nch; ++k) {
        const uint32_t b = nd.first + k * kChunk;
        chunks[c0 + k] = make_uint4(s, b, b + kChunk < end ? b + kChunk : end, 0u);
    }
}

// ---- one level: bins, split, partition ----------------------------------------------

__global__ void k_bins_clear(uint32_t* bins, uint32_t words)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += gridDim.x * blockDim.x) {
        const uint32_t f = i % 7u;
        bins[i] = f == 0 ? 0u : (f < 4 ? KEY_FMAX : KEY_NFMAX);
    }
}

// bins[slot][axis][bin][7]: count, vertex min keys (3), vertex max keys (3)
__global__ __launch_bounds__(kWG) void k_bins(const uint4* __restrict__ chunks, uint32_t nchunks, const uint32_t* __restrict__ active,
                                              const BNode* __restrict__ nodes, const float4* __restrict__ rc,
                                              const float4* __restrict__ rl, const float4* __restrict__ rh, uint32_t B,
                                              uint32_t* bins)
{
    __shared__ uint32_t sb[kWG / 64][3 * kDevBuildMaxBins * 7];
    const uint32_t W = 3u * B * 7u;
    const uint32_t wave = threadIdx.x >> 6;
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const uint4 ch = chunks[c];
        const BNode& nd = nodes[active[ch.x]];
        float lo[3], scale[3];
        bool live[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = kfloat(nd.cmn[a]);
            const float hi = kfloat(nd.cmx[a]);
            live[a] = lo[a] != hi;
            scale[a] = live[a] ? (float)B / (hi - lo[a]) : 0.0f;
        }
        for (uint32_t e = threadIdx.x; e < (kWG / 64) * W; e += kWG) {
            const uint32_t i = e % W, f = i % 7u;
            sb[e / W][i] = f == 0 ? 0u : (f < 4 ? KEY_FMAX : KEY_NFMAX);
        }
        __syncthreads();
        for (uint32_t p = ch.y + threadIdx.x; p < ch.z; p += kWG) {
            const float4 c4 = rc[p], l4 = rl[p], h4 = rh[p];
            const uint32_t kl0 = fkey(l4.x), kl1 = fkey(l4.y), kl2 = fkey(l4.z);
            const uint32_t kh0 = fkey(h4.x), kh1 = fkey(h4.y), kh2 = fkey(h4.z);
            for (int a = 0; a < 3; ++a) {
                if (!live[a]) continue;
                const uint32_t b = bin_of(rec_get(c4, a), lo[a], scale[a], B);
                uint32_t* q = &sb[wave][((uint32_t)a * B + b) * 7u];
                atomicAdd(q, 1u);
                atomicMin(q + 1, kl0); atomicMin(q + 2, kl1); atomicMin(q + 3, kl2);
                atomicMax(q + 4, kh0); atomicMax(q + 5, kh1); atomicMax(q + 6, kh2);
            }
        }
        __syncthreads();
        uint32_t* gb = bins + (size_t)ch.x * W;
        for (uint32_t i = threadIdx.x; i < W; i += kWG) {
            const uint32_t f = i % 7u, base = i - f;
            uint32_t n_in = 0;
            for (uint32_t w = 0; w < kWG / 64; ++w) n_in += sb[w][base];
            if (n_in == 0) continue;
            uint32_t v = sb[0][i];
            for (uint32_t w = 1; w < kWG / 64; ++w) {
                const uint32_t u = sb[w][i];
                v = f == 0 ? v + u : (f < 4 ? (u < v ? u : v) : (u > v ? u : v));
            }
            if (f == 0) atomicAdd(gb + i, v);
            else if (f < 4) atomicMin(gb + i, v);
            else atomicMax(gb + i, v);
        }
        __syncthreads();
    }
}

// find_best_split_plane (bvh_builder.cpp:170-250) and the decisions of
// subdivide_node (bvh_builder.cpp:100-168) for every node of the level.
__global__ void k_split(uint32_t a, const uint32_t* __restrict__ active, BNode* nodes, const uint32_t* __restrict__ bins, uint32_t B,
                        uint32_t single_prim, SplitInfo* info, uint32_t* next_list, uint32_t* small_list, Counters* ctr,
                        uint32_t out_buf)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a) return;
    const uint32_t id = active[s];
    BNode& nd = nodes[id];
    const uint32_t* bn = bins + (size_t)s * 3u * B * 7u;
    float best = FLT_MAX;
    uint32_t baxis = 0, bsplit = 0;
    float ra[kDevBuildMaxBins];
    for (uint32_t axis = 0; axis < 3; ++axis) {
        const float lo = kfloat(nd.cmn[axis]), hi = kfloat(nd.cmx[axis]);
        if (lo == hi) continue;
        const uint32_t* ab = bn + axis * B * 7u;
        float rmn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, rmx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        uint32_t rs = 0;
        for (uint32_t i = 0; i + 1 < B; ++i) {
            const uint32_t j = B - 1 - i;
            rs += ab[j * 7u];
            for (int k = 0; k < 3; ++k) {
                rmn[k] = gmin(rmn[k], kfloat(ab[j * 7u + 1 + k]));
                rmx[k] = gmax(rmx[k], kfloat(ab[j * 7u + 4 + k]));
            }
            ra[B - 2 - i] = (float)rs * half_area(rmx[0] - rmn[0], rmx[1] - rmn[1], rmx[2] - rmn[2]);
        }
        float lmn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, lmx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        uint32_t ls = 0;
        for (uint32_t i = 0; i + 1 < B; ++i) {
            ls += ab[i * 7u];
            for (int k = 0; k < 3; ++k) {
                lmn[k] = gmin(lmn[k], kfloat(ab[i * 7u + 1 + k]));
                lmx[k] = gmax(lmx[k], kfloat(ab[i * 7u + 4 + k]));
            }
            const float la = (float)ls * half_area(lmx[0] - lmn[0], lmx[1] - lmn[1], lmx[2] - lmn[2]);
            const float cost = la + ra[i];
            if (cost < best) { best = cost; baxis = axis; bsplit = i + 1; }
        }
    }
    SplitInfo si{};
    si.node = id;
    bool leaf;
    if (single_prim) {
        leaf = nd.count == 1;
    } else {
        const float parent = (float)nd.count * half_area(nd.mx[0] - nd.mn[0], nd.mx[1] - nd.mn[1], nd.mx[2] - nd.mn[2]);
        leaf = best >= parent;
    }
    if (!leaf && best == FLT_MAX) leaf = true;
    if (leaf) {
        si.mode = 0;
        nd.kind = kLeaf;
        nd.desc = 0;
        atomicMax(&ctr->max_leaf, nd.count);
        info[s] = si;
        return;
    }
    const float lo = kfloat(nd.cmn[baxis]), hi = kfloat(nd.cmx[baxis]);
    const uint32_t* ab = bn + baxis * B * 7u;
    uint32_t L = 0;
    for (uint32_t b = 0; b < bsplit; ++b) L += ab[b * 7u];
    si.axis = baxis;
    si.split = bsplit;
    si.L = L;
    si.lo = lo;
    si.scale = (float)B / (hi - lo);
    if (L == 0 || L == nd.count) {
        // the swap loop still ran over the range; the node stays a leaf
        si.mode = 2;
        nd.kind = kLeaf;
        nd.desc = 0;
        atomicMax(&ctr->max_leaf, nd.count);
        info[s] = si;
        return;
    }
    si.mode = 1;
    const uint32_t l = atomicAdd(&ctr->nodes, 2u), r = l + 1;
    for (uint32_t side = 0; side < 2; ++side) {
        BNode& c = nodes[l + side];
        float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        const uint32_t b0 = side ? bsplit : 0, b1 = side ? B : bsplit;
        for (uint32_t b = b0; b < b1; ++b) {
            if (ab[b * 7u] == 0) continue;
            for (int k = 0; k < 3; ++k) {
                mn[k] = gmin(mn[k], kfloat(ab[b * 7u + 1 + k]));
                mx[k] = gmax(mx[k], kfloat(ab[b * 7u + 4 + k]));
            }
        }
        for (int k = 0; k < 3; ++k) {
            c.mn[k] = mn[k];
            c.mx[k] = mx[k];
            c.cmn[k] = KEY_FMAX;
            c.cmx[k] = KEY_NFMAX;
        }
        for (int k = 0; k < 6; ++k) c.fz[k] = kNone;
        c.first = side ? nd.first + L : nd.first;
        c.count = side ? nd.count - L : L;
        c.left = c.right = kNone;
        c.desc = 0;
        c.fid = 0;
        c.cursor = 0;
        c.block = 0;
        c.inner = 0;
        c.buf = out_buf;
        if (c.count <= kSmall) {
            c.kind = kSmallRoot;
            small_list[atomicAdd(&ctr->small, 1u)] = l + side;
        } else {
            c.kind = kOpen;
            next_list[atomicAdd(&ctr->next, 1u)] = l + side;
        }
    }
    nd.kind = kInner;
    nd.left = l;
    nd.right = r;
    info[s] = si;
}


// Per chunk: the number of right elements in [0, L) and of left elements in
// [L, n); the centroid bounds of children that stay in the level passes.
__global__ __launch_bounds__(kWG) void k_part_count(const uint4* __restrict__ chunks, uint32_t nchunks, const SplitInfo* __restrict__ info,
                                                    BNode* nodes, const float4* __restrict__ rc, uint32_t B, uint2* cnt)
{
    __shared__ uint32_t s_w[2 * (kWG / 64)];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const uint4 ch = chunks[c];
        const SplitInfo si = info[ch.x];
        if (si.mode == 0) {
            if (threadIdx.x == 0) cnt[c] = make_uint2(0u, 0u);
            continue;
        }
        const BNode& nd = nodes[si.node];
        const uint32_t first = nd.first;
        bool want[2] = {false, false};
        uint32_t child[2] = {kNone, kNone};
        if (si.mode == 1) {
            child[0] = nd.left;
            child[1] = nd.right;
            want[0] = nodes[child[0]].kind == kOpen;
            want[1] = nodes[child[1]].kind == kOpen;
        }
        uint32_t nfr = 0, nbl = 0;
        uint32_t kmn[2][3], kmx[2][3];
        for (int s = 0; s < 2; ++s)
            for (int k = 0; k < 3; ++k) { kmn[s][k] = KEY_FMAX; kmx[s][k] = KEY_NFMAX; }
        for (uint32_t p = ch.y + threadIdx.x; p < ch.z; p += kWG) {
            const float4 c4 = rc[p];
            const bool left = bin_of(rec_get(c4, (int)si.axis), si.lo, si.scale, B) < si.split;
            const uint32_t rel = p - first;
            nfr += (rel < si.L && !left) ? 1u : 0u;
            nbl += (rel >= si.L && left) ? 1u : 0u;
            const int s = left ? 0 : 1;
            if (want[s]) {
                for (int k = 0; k < 3; ++k) {
                    const uint32_t key = fkey(rec_get(c4, k));
                    kmn[s][k] = key < kmn[s][k] ? key : kmn[s][k];
                    kmx[s][k] = key > kmx[s][k] ? key : kmx[s][k];
                }
            }
        }
        nfr = wave_sum_u(nfr);
        nbl = wave_sum_u(nbl);
        for (int s = 0; s < 2; ++s) {
            if (!want[s]) continue;
            for (int k = 0; k < 3; ++k) {
                const uint32_t a = wave_min_u(kmn[s][k]), b = wave_max_u(kmx[s][k]);
                if (lane == 0) {
                    if (a != KEY_FMAX) atomicMin(&nodes[child[s]].cmn[k], a);
                    if (b != KEY_NFMAX) atomicMax(&nodes[child[s]].cmx[k], b);
                }
            }
        }
        if (lane == 0) { s_w[2 * wave] = nfr; s_w[2 * wave + 1] = nbl; }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t a = 0, b = 0;
            for (uint32_t w = 0; w < kWG / 64; ++w) { a += s_w[2 * w]; b += s_w[2 * w + 1]; }
            cnt[c] = make_uint2(a, b);
        }
        __syncthreads();
    }
}

// The partition of every split node of the level (closed form of the swap
// loop, see the top of this file).  TABLES: the positions of the right
// elements in [0, L) by rank and of the left elements in [L, n) by rank from
// the back.  MOVE: every element to its place in the other record buffer (for
// a node that stays a leaf, its index entry to its place in the output).
template <bool MOVE>
__global__ __launch_bounds__(kWG) void k_part(const uint4* __restrict__ chunks, uint32_t nchunks, const SplitInfo* __restrict__ info,
                                              const BNode* __restrict__ nodes, const uint2* __restrict__ off, const uint2* __restrict__ slot_ch,
                                              uint32_t B, const float4* __restrict__ rc, const float4* __restrict__ rl,
                                              const float4* __restrict__ rh, float4* oc, float4* ol, float4* oh, uint32_t* tfr,
                                              uint32_t* tbl, uint32_t* idx_out, Counters* ctr)
{
    __shared__ uint32_t s_w[2 * (kWG / 64)];
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const uint4 ch = chunks[c];
        const SplitInfo si = info[ch.x];
        if (si.mode == 0) {
            if (MOVE)
                for (uint32_t p = ch.y + threadIdx.x; p < ch.z; p += kWG) idx_out[p] = __float_as_uint(rc[p].w);
            continue;
        }
        const BNode& nd = nodes[si.node];
        const uint32_t first = nd.first, n = nd.count, L = si.L;
        const uint2 sc = slot_ch[ch.x];
        const uint32_t rf = off[sc.x + sc.y].x - off[sc.x].x;
        const uint32_t base_fr = off[c].x - off[sc.x].x, base_bl = off[c].y - off[sc.x].y;
        uint32_t run_fr = 0, run_bl = 0;
        for (uint32_t row = 0; row < kRows; ++row) {
            const uint32_t p = ch.y + row * kWG + threadIdx.x;
            const bool valid = p < ch.z;
            const uint32_t rel = p - first;
            float4 c4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            bool left = false;
            if (valid) {
                c4 = rc[p];
                left = bin_of(rec_get(c4, (int)si.axis), si.lo, si.scale, B) < si.split;
            }
            const bool fr = valid && rel < L && !left;
            const bool bl = valid && rel >= L && left;
            uint32_t ex_fr, ex_bl, t_fr, t_bl;
            block_scan2(fr, bl, ex_fr, ex_bl, t_fr, t_bl, s_w);
            const uint32_t m_fr = base_fr + run_fr + ex_fr + 1u;       // rank from the front, 1-based
            const uint32_t m_bl = rf - (base_bl + run_bl + ex_bl);     // rank from the back, 1-based
            if (!MOVE) {
                if ((fr && (m_fr < 1u || m_fr > n)) || (bl && (m_bl < 1u || m_bl > n))) { atomicOr(&ctr->bad_index, 1u); continue; }
                if (fr) tfr[first + m_fr - 1u] = rel;
                if (bl) tbl[first + m_bl - 1u] = rel;
            } else if (valid) {
                if ((fr && (m_fr < 1u || m_fr > n)) || (bl && (m_bl < 1u || m_bl > n)) || rf > n) {
                    atomicOr(&ctr->bad_index, 1u);
                    continue;
                }
                uint32_t dest;
                if (rel < L && left) dest = rel;
                else if (bl) dest = tfr[first + m_bl - 1u];
                else if (fr) dest = m_fr == 1u ? n - 1u : tbl[first + m_fr - 2u] - 1u;
                else dest = rel > L ? rel - 1u : (rf == 0 ? n - 1u : tbl[first + rf - 1u] - 1u);
                if (dest >= n) { atomicOr(&ctr->bad_index, 1u); continue; }
                if (si.mode == 2) {
                    idx_out[first + dest] = __float_as_uint(c4.w);
                } else {
                    oc[first + dest] = c4;
                    ol[first + dest] = rl[p];
                    oh[first + dest] = rh[p];
                }
            }
            run_fr += t_fr;
            run_bl += t_bl;
        }
    }
}

// Signed-zero fix-up of node bounds: the first position (in the order the
// node's range holds) whose bound component is a zero equal to the node's.
__global__ __launch_bounds__(kWG) void k_zero_scan(const uint4* __restrict__ chunks, uint32_t nchunks, const SplitInfo* __restrict__ info,
                                                   BNode* nodes, const float4* __restrict__ rl, const float4* __restrict__ rh,
                                                   uint32_t root_mode)
{
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const uint4 ch = chunks[c];
        uint32_t first = 0, L = 0, l = 0, r = 0;
        if (!root_mode) {
            const SplitInfo si = info[ch.x];
            if (si.mode != 1) continue;
            const BNode& nd = nodes[si.node];
            first = nd.first;
            L = si.L;
            l = nd.left;
            r = nd.right;
        }
        for (uint32_t p = ch.y + threadIdx.x; p < ch.z; p += kWG) {
            const uint32_t id = root_mode ? 0u : ((p - first) < L ? l : r);
            BNode& t = nodes[id];
            if (!root_mode && t.kind != kOpen) continue;
            const float4 l4 = rl[p], h4 = rh[p];
            for (int k = 0; k < 3; ++k) {
                if (t.mn[k] == 0.0f && rec_get(l4, k) == 0.0f) atomicMin(&t.fz[k], p);
                if (t.mx[k] == 0.0f && rec_get(h4, k) == 0.0f) atomicMin(&t.fz[3 + k], p);
            }
        }
    }
}

__global__ void k_zero_resolve(const uint32_t* list, uint32_t cnt, BNode* nodes, const float4* rl, const float4* rh)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    BNode& t = nodes[list[i]];
    for (int k = 0; k < 3; ++k) {
        if (t.fz[k] != kNone) t.mn[k] = rec_get(rl[t.fz[k]], k);
        if (t.fz[3 + k] != kNone) t.mx[k] = rec_get(rh[t.fz[3 + k]], k);
    }
}

// ---- subtrees of at most kSmall triangles: one wave each ---------------------------

// The reference's recursion (subdivide_node, find_best_split_plane,
// calc_node_min_max) over one small node's range, one triangle per lane.
// Nodes go to a block with subtree-local numbering: the root's children are 0
// and 1, later pairs follow in the reference's depth-first order.
__global__ __launch_bounds__(kWG) void k_wave_build(const uint32_t* __restrict__ small_list, uint32_t ns, BNode* nodes,
                                                    const float4* __restrict__ rc0, const float4* __restrict__ rl0,
                                                    const float4* __restrict__ rh0, const float4* __restrict__ rc1,
                                                    const float4* __restrict__ rl1, const float4* __restrict__ rh1, uint32_t B,
                                                    uint32_t single_prim, wf_bvh_node* blocks, uint32_t* idx_out, Counters* ctr)
{
    constexpr uint32_t NW = kWG / 64;
    __shared__ uint32_t s_bins[NW][kDevBuildMaxBins * 7];
    __shared__ uint32_t s_tfr[NW][64], s_tbl[NW][64];
    __shared__ float4 s_rec[NW][3][64];
    __shared__ uint32_t s_stk[NW][3 * (kSmall + 2)];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (uint32_t w = blockIdx.x * NW + wave; w < ns; w += gridDim.x * NW) {
        const uint32_t id = small_list[w];
        BNode& S = nodes[id];
        const uint32_t F = S.first, c = S.count;
        const bool b1 = S.buf != 0;
        float4 vc = make_float4(0.0f, 0.0f, 0.0f, 0.0f), vl = vc, vh = vc;
        if (lane < c) {
            vc = (b1 ? rc1 : rc0)[F + lane];
            vl = (b1 ? rl1 : rl0)[F + lane];
            vh = (b1 ? rh1 : rh0)[F + lane];
        }
        uint32_t block = 0;
        if (lane == 0) block = atomicAdd(&ctr->blocks, 2u * (c - 1u));
        block = uni(block);
        uint32_t cursor = 0, sp = 1, max_leaf = 0;
        if (lane == 0) { s_stk[wave][0] = 0u; s_stk[wave][1] = c; s_stk[wave][2] = kNone; }
        wave_sync();
        while (sp > 0) {
            --sp;
            const uint32_t lo = uni(s_stk[wave][3 * sp]), cnt = uni(s_stk[wave][3 * sp + 1]), lid = uni(s_stk[wave][3 * sp + 2]);
            const bool act = lane >= lo && lane < lo + cnt;
            // calc_node_min_max (bvh_builder.cpp:71-93)
            float vmn[3], vmx[3], cmn[3], cmx[3];
            for (int k = 0; k < 3; ++k) {
                const float x = rec_get(vl, k), y = rec_get(vh, k), z = rec_get(vc, k);
                vmn[k] = kfloat(wave_min_u(act ? fkey(x) : KEY_FMAX));
                if (vmn[k] == 0.0f) {
                    const uint64_t m = __ballot(act && x == 0.0f);
                    vmn[k] = __shfl(x, __ffsll((unsigned long long)m) - 1, 64);
                }
                vmx[k] = kfloat(wave_max_u(act ? fkey(y) : KEY_NFMAX));
                if (vmx[k] == 0.0f) {
                    const uint64_t m = __ballot(act && y == 0.0f);
                    vmx[k] = __shfl(y, __ffsll((unsigned long long)m) - 1, 64);
                }
                cmn[k] = kfloat(wave_min_u(act ? fkey(z) : KEY_FMAX));
                cmx[k] = kfloat(wave_max_u(act ? fkey(z) : KEY_NFMAX));
            }
            // find_best_split_plane (bvh_builder.cpp:170-250): bins in LDS, the
            // prefix and suffix sweeps across lanes (lane i = plane after bin i)
            float best = FLT_MAX;
            uint32_t baxis = 0, bsplit = 0;
            for (uint32_t axis = 0; axis < 3; ++axis) {
                const float lo_c = cmn[axis], hi_c = cmx[axis];
                if (lo_c == hi_c) continue;
                const float scale = (float)B / (hi_c - lo_c);
                for (uint32_t e = lane; e < B * 7u; e += 64u) {
                    const uint32_t f = e % 7u;
                    s_bins[wave][e] = f == 0 ? 0u : (f < 4 ? KEY_FMAX : KEY_NFMAX);
                }
                wave_sync();
                if (act) {
                    const uint32_t b = bin_of(rec_get(vc, (int)axis), lo_c, scale, B);
                    uint32_t* q = &s_bins[wave][b * 7u];
                    atomicAdd(q, 1u);
                    atomicMin(q + 1, fkey(vl.x)); atomicMin(q + 2, fkey(vl.y)); atomicMin(q + 3, fkey(vl.z));
                    atomicMax(q + 4, fkey(vh.x)); atomicMax(q + 5, fkey(vh.y)); atomicMax(q + 6, fkey(vh.z));
                }
                wave_sync();
                // lane j: bin j (prefix side) and bin B-1-j (suffix side), as
                // order keys; inclusive scans give the left unions of the
                // planes and, reversed, the right ones
                uint32_t pc = 0, rc = 0;
                uint32_t pk[6] = {KEY_FMAX, KEY_FMAX, KEY_FMAX, KEY_NFMAX, KEY_NFMAX, KEY_NFMAX};
                uint32_t rk[6] = {KEY_FMAX, KEY_FMAX, KEY_FMAX, KEY_NFMAX, KEY_NFMAX, KEY_NFMAX};
                if (lane < B) {
                    const uint32_t* bp = &s_bins[wave][lane * 7u];
                    const uint32_t* bq = &s_bins[wave][(B - 1u - lane) * 7u];
                    pc = bp[0];
                    rc = bq[0];
                    for (int k = 0; k < 6; ++k) { pk[k] = bp[1 + k]; rk[k] = bq[1 + k]; }
                }
                wave_sync();
                pc = __ockl_wfscan_add_u32(pc, true);
                rc = __ockl_wfscan_add_u32(rc, true);
                for (int k = 0; k < 3; ++k) {
                    pk[k] = __ockl_wfscan_min_u32(pk[k], true);
                    pk[3 + k] = __ockl_wfscan_max_u32(pk[3 + k], true);
                    rk[k] = __ockl_wfscan_min_u32(rk[k], true);
                    rk[3 + k] = __ockl_wfscan_max_u32(rk[3 + k], true);
                }
                // right side of plane `lane` (bins lane+1 .. B-1) = reversed scan at lane B-2-lane
                const int src = (int)B - 2 - (int)lane;
                const int srcl = src >= 0 ? src : 0;
                const uint32_t rs = __shfl(rc, srcl, 64);
                float rmn[3], rmx[3], pmn[3], pmx[3];
                for (int k = 0; k < 3; ++k) {
                    rmn[k] = kfloat(__shfl(rk[k], srcl, 64));
                    rmx[k] = kfloat(__shfl(rk[3 + k], srcl, 64));
                    pmn[k] = kfloat(pk[k]);
                    pmx[k] = kfloat(pk[3 + k]);
                }
                // the first plane of least cost (the reference's strict '<'
                // scan): costs are >= 0, so (cost key, plane) orders as the
                // scan does once -0 is made +0
                uint64_t key = ~0ull;
                if (lane + 1u < B) {
                    const float la = (float)pc * half_area(pmx[0] - pmn[0], pmx[1] - pmn[1], pmx[2] - pmn[2]);
                    const float ra = (float)rs * half_area(rmx[0] - rmn[0], rmx[1] - rmn[1], rmx[2] - rmn[2]);
                    const float cst = la + ra;
                    if (cst < FLT_MAX) key = ((uint64_t)fkey(cst + 0.0f) << 32) | lane;
                }
                key = __ockl_wfred_min_u64(key);
                if (key != ~0ull) {
                    const float cost = kfloat((uint32_t)(key >> 32));
                    if (cost < best) { best = cost; baxis = axis; bsplit = (uint32_t)key + 1u; }
                }
            }
            // subdivide_node's decisions (bvh_builder.cpp:105-117)
            bool leaf;
            if (single_prim) {
                leaf = cnt == 1u;
            } else {
                const float parent = (float)cnt * half_area(vmx[0] - vmn[0], vmx[1] - vmn[1], vmx[2] - vmn[2]);
                leaf = best >= parent;
            }
            if (!leaf && best == FLT_MAX) leaf = true;
            uint32_t L = 0;
            if (!leaf) {
                // the swap partition (bvh_builder.cpp:123-137) in closed form
                const float lo_c = cmn[baxis];
                const float scale = (float)B / (cmx[baxis] - lo_c);
                const bool left = act && bin_of(rec_get(vc, (int)baxis), lo_c, scale, B) < bsplit;
                L = (uint32_t)__popcll(__ballot(left));
                const uint32_t rel = lane - lo;
                const bool fr = act && rel < L && !left;
                const bool bl = act && rel >= L && left;
                const uint64_t mfr = __ballot(fr), mbl = __ballot(bl);
                const uint32_t rf = (uint32_t)__popcll(mfr);
                const uint32_t m_fr = lanes_below(mfr, lane) + 1u;
                const uint32_t m_bl = (uint32_t)__popcll(mbl >> lane);
                if (fr) s_tfr[wave][m_fr - 1u] = rel;
                if (bl) s_tbl[wave][m_bl - 1u] = rel;
                wave_sync();
                if (act) {
                    uint32_t dest;
                    if (rel < L && left) dest = rel;
                    else if (bl) dest = s_tfr[wave][m_bl - 1u];
                    else if (fr) dest = m_fr == 1u ? cnt - 1u : s_tbl[wave][m_fr - 2u] - 1u;
                    else dest = rel > L ? rel - 1u : (rf == 0 ? cnt - 1u : s_tbl[wave][rf - 1u] - 1u);
                    if (dest >= cnt) { atomicOr(&ctr->bad_index, 1u); dest = rel; }
                    s_rec[wave][0][lo + dest] = vc;
                    s_rec[wave][1][lo + dest] = vl;
                    s_rec[wave][2][lo + dest] = vh;
                }
                wave_sync();
                if (act) {
                    vc = s_rec[wave][0][lane];
                    vl = s_rec[wave][1][lane];
                    vh = s_rec[wave][2][lane];
                }
                wave_sync();
                if (L == 0 || L == cnt) leaf = true;
            }
            wf_bvh_node o;
            for (int k = 0; k < 3; ++k) { o.aabb_min[k] = vmn[k]; o.aabb_max[k] = vmx[k]; }
            if (leaf) {
                o.left_first = F + lo;
                o.prim_count = cnt;
                max_leaf = cnt > max_leaf ? cnt : max_leaf;
            } else {
                const uint32_t lc = cursor;
                cursor += 2u;
                o.left_first = lc;
                o.prim_count = 0;
                if (sp + 2u > kSmall + 2u) { if (lane == 0) atomicOr(&ctr->bad_index, 1u); break; }
                if (lane == 0) {
                    s_stk[wave][3 * sp] = lo + L;
                    s_stk[wave][3 * sp + 1] = cnt - L;
                    s_stk[wave][3 * sp + 2] = lc + 1u;
                    s_stk[wave][3 * sp + 3] = lo;
                    s_stk[wave][3 * sp + 4] = L;
                    s_stk[wave][3 * sp + 5] = lc;
                }
                sp += 2u;
                wave_sync();
            }
            if (lane == 0) {
                if (lid == kNone) {
                    for (int k = 0; k < 3; ++k) { S.mn[k] = vmn[k]; S.mx[k] = vmx[k]; }
                    S.inner = leaf ? 0u : 1u;
                } else {
                    blocks[block + lid] = o;
                }
            }
        }
        if (lane < c) idx_out[F + lane] = __float_as_uint(vc.w);
        if (lane == 0) {
            S.desc = cursor;
            S.block = block;
            atomicMax(&ctr->max_leaf, max_leaf);
        }
    }
}

// ---- renumbering in the reference's allocation order ------------------------------

__global__ void k_desc(const uint32_t* list, uint32_t cnt, BNode* nodes)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    BNode& x = nodes[list[i]];
    if (x.kind == kInner) x.desc = 2u + nodes[x.left].desc + nodes[x.right].desc;
}

__global__ void k_assign(const uint32_t* list, uint32_t cnt, BNode* nodes)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    const BNode& x = nodes[list[i]];
    if (x.kind != kInner) return;
    BNode& l = nodes[x.left];
    BNode& r = nodes[x.right];
    l.fid = x.cursor;
    r.fid = x.cursor + 1u;
    l.cursor = x.cursor + 2u;
    r.cursor = x.cursor + 2u + l.desc;
}

__global__ void k_emit_upper(const uint32_t* list, uint32_t cnt, const BNode* nodes, wf_bvh_node* out, uint32_t cap, Counters* ctr)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    const BNode& x = nodes[list[i]];
    if (x.fid >= cap) { atomicOr(&ctr->bad_index, 1u); return; }
    const bool inner = x.kind == kInner;
    wf_bvh_node o;
    for (int k = 0; k < 3; ++k) { o.aabb_min[k] = x.mn[k]; o.aabb_max[k] = x.mx[k]; }
    o.left_first = inner ? x.cursor : x.first;
    o.prim_count = inner ? 0u : x.count;
    out[x.fid] = o;
}

__global__ __launch_bounds__(kWG) void k_emit_small(const uint32_t* small_list, uint32_t ns, const BNode* nodes,
                                                    const wf_bvh_node* blocks, wf_bvh_node* out, uint32_t cap, Counters* ctr)
{
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t w = blockIdx.x * (kWG / 64) + (threadIdx.x >> 6); w < ns; w += gridDim.x * (kWG / 64)) {
        const BNode& S = nodes[small_list[w]];
        if (S.fid >= cap || S.cursor + S.desc > cap) {
            if (lane == 0) atomicOr(&ctr->bad_index, 1u);
            continue;
        }
        if (lane == 0) {
            wf_bvh_node o;
            for (int k = 0; k < 3; ++k) { o.aabb_min[k] = S.mn[k]; o.aabb_max[k] = S.mx[k]; }
            o.left_first = S.inner ? S.cursor : S.first;
            o.prim_count = S.inner ? 0u : S.count;
            out[S.fid] = o;
        }
        for (uint32_t k = lane; k < S.desc; k += 64u) {
            wf_bvh_node b = blocks[S.block + k];
            if (b.prim_count == 0) b.left_first += S.cursor;
            out[S.cursor + k] = b;
        }
    }
}

__global__ void k_tris_from_index(const float* __restrict__ pos, const uint32_t* __restrict__ idx, uint32_t n, GpuTri* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t t = idx[i];
    const float* q = pos + 9ull * t;
    GpuTri g;
    for (int k = 0; k < 3; ++k) {
        g.v0[k] = q[k];
        g.e1[k] = q[3 + k] - q[k];
        g.e2[k] = q[6 + k] - q[k];
    }
    g.prim = t;
    g.pad0 = g.pad1 = 0u;
    out[i] = g;
}

inline uint32_t grid_for(uint64_t items, uint32_t per_block, uint32_t cap)
{
    const uint64_t g = (items + per_block - 1) / per_block;
    return (uint32_t)(g == 0 ? 1 : (g < cap ? g : cap));
}

} // namespace

void DevBuildScratch::release()
{
    if (fixed) return;   // the memory is the caller's
    for (int b = 0; b < 2; ++b) { rc[b].release(); rl[b].release(); rh[b].release(); }
    for (DevBuf* d : {&tfr, &tbl, &nodes, &lists, &small, &blocks, &info, &bins, &chunks, &cnt, &off, &slot_ch, &ctr}) d->release();
}

static DevBuf* scratch_buffers(DevBuildScratch& S, DevBuf* out[kDevBuildBuffers])
{
    DevBuf* list[kDevBuildBuffers] = {&S.rc[0], &S.rl[0], &S.rh[0], &S.rc[1], &S.rl[1], &S.rh[1], &S.tfr, &S.tbl, &S.nodes,
                                      &S.lists, &S.small, &S.blocks, &S.info, &S.bins, &S.chunks, &S.cnt, &S.off,
                                      &S.slot_ch, &S.ctr};
    for (int k = 0; k < kDevBuildBuffers; ++k) out[k] = list[k];
    return out[0];
}

void device_build_sizes(uint32_t n, uint32_t B, size_t out[kDevBuildBuffers])
{
    const size_t maxa = n / (kSmall + 1) + 2;
    const size_t maxch = n / kChunk + maxa + 2;
    const size_t scan_n = (maxch > maxa ? maxch : maxa) + 1;
    for (int k = 0; k < 6; ++k) out[k] = 16ull * n;
    out[6] = 4ull * n;                               // tfr
    out[7] = 4ull * n;                               // tbl
    out[8] = sizeof(BNode) * (2ull * n + 1);         // nodes
    out[9] = 4ull * (2ull * n + 1);                  // lists
    out[10] = 4ull * n;                              // small
    out[11] = sizeof(wf_bvh_node) * 2ull * n;        // blocks
    out[12] = sizeof(SplitInfo) * maxa;              // info
    out[13] = 4ull * 3 * B * 7 * maxa;               // bins
    out[14] = 16ull * maxch;                         // chunks
    out[15] = 8ull * scan_n;                         // cnt
    out[16] = 8ull * (scan_n + 1);                   // off
    out[17] = 8ull * maxa;                           // slot_ch
    out[18] = sizeof(Counters);                      // ctr
}

hipError_t launch_tris_from_index(const float* d_pos, const uint32_t* d_idx, uint32_t n, GpuTri* d_tris, hipStream_t stream)
{
    k_tris_from_index<<<grid_for(n, 256, 0x7FFFFFFF), 256, 0, stream>>>(d_pos, d_idx, n, d_tris);
    return hipGetLastError();
}

wf_status device_build_blas(const float* d_pos, uint32_t n, const wf_build_options& opt, wf_bvh_node* d_nodes,
                            uint32_t* d_idx, DevBuildScratch& S, hipStream_t st, DevBuildResult& res, std::string& err)
{
    if (!d_pos || !d_nodes || !d_idx || n == 0) { err = "null pointer or empty mesh"; return WF_ERROR_INVALID_ARGUMENT; }
    const uint32_t B = opt.interval_count;
    if (B < 2 || B > kDevBuildMaxBins) { err = "the device builder takes 2..64 intervals"; return WF_ERROR_INVALID_ARGUMENT; }
    if (n > (1u << 28)) { err = "mesh too large for the device builder"; return WF_ERROR_CAPACITY; }
    const uint32_t single = opt.subdivide_single_prim ? 1u : 0u;
    {
        size_t sizes[kDevBuildBuffers];
        DevBuf* bufs[kDevBuildBuffers];
        device_build_sizes(n, B, sizes);
        scratch_buffers(S, bufs);
        for (int k = 0; k < kDevBuildBuffers; ++k) {
            if (S.fixed) {
                if (!bufs[k]->p || bufs[k]->bytes < sizes[k]) { err = "mesh too large for the fixed scratch"; return WF_ERROR_CAPACITY; }
            } else if (bufs[k]->ensure(sizes[k]) != hipSuccess) {
                err = "device builder scratch allocation failed";
                return WF_ERROR_OUT_OF_MEMORY;
            }
        }
    }
    hipError_t e = hipSuccess;

    float4* rc[2] = {(float4*)S.rc[0].p, (float4*)S.rc[1].p};
    float4* rl[2] = {(float4*)S.rl[0].p, (float4*)S.rl[1].p};
    float4* rh[2] = {(float4*)S.rh[0].p, (float4*)S.rh[1].p};
    uint32_t* tfr = (uint32_t*)S.tfr.p;
    uint32_t* tbl = (uint32_t*)S.tbl.p;
    BNode* nodes = (BNode*)S.nodes.p;
    uint32_t* lists = (uint32_t*)S.lists.p;
    uint32_t* small = (uint32_t*)S.small.p;
    wf_bvh_node* blocks = (wf_bvh_node*)S.blocks.p;
    SplitInfo* info = (SplitInfo*)S.info.p;
    uint32_t* bins = (uint32_t*)S.bins.p;
    uint4* chunks = (uint4*)S.chunks.p;
    uint2* cnt = (uint2*)S.cnt.p;
    uint2* off = (uint2*)S.off.p;
    uint2* slot_ch = (uint2*)S.slot_ch.p;
    Counters* ctr = (Counters*)S.ctr.p;
    Counters hc{};

    auto check = [&](const char* what) -> bool {
        if (e == hipSuccess) e = hipGetLastError();
        if (e != hipSuccess) { err = std::string(what) + ": " + hipGetErrorString(e); return false; }
        return true;
    };
    auto read_ctr = [&]() -> bool {
        if (e == hipSuccess) e = hipMemcpyAsync(&hc, ctr, sizeof(Counters), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        return check("device build counters");
    };
    uint32_t nch = 0;
    // chunk list of the active nodes list[0..a)
    auto make_chunks = [&](const uint32_t* list, uint32_t a) -> bool {
        k_chunk_count<<<grid_for(a, 256, 1u << 20), 256, 0, st>>>(list, a, nodes, cnt);
        k_scan<<<1, 1024, 0, st>>>(cnt, a, off);
        k_chunk_fill<<<grid_for(a, 256, 1u << 20), 256, 0, st>>>(list, a, nodes, off, chunks, slot_ch);
        uint2 tot{};
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(&tot, off + a, sizeof(uint2), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        nch = tot.x;
        return check("device build chunk list");
    };

    k_ctr_init<<<1, 64, 0, st>>>(ctr);
    k_prep<<<grid_for(n, 256, 2048), 256, 0, st>>>(d_pos, n, rc[0], rl[0], rh[0], ctr);
    if (!read_ctr()) return WF_ERROR_DEVICE;
    if (hc.nonfinite) { err = "non-finite vertex position"; return WF_ERROR_INVALID_ARGUMENT; }
    const bool negzero = hc.negzero != 0;
    k_root_init<<<1, 1, 0, st>>>(nodes, n, ctr, lists, small);
    uint32_t a = n > kSmall ? 1u : 0u;
    if (a) {
        if (!make_chunks(lists, a)) return WF_ERROR_DEVICE;
        if (negzero) {
            k_zero_scan<<<grid_for(nch, 1, 8192), kWG, 0, st>>>(chunks, nch, info, nodes, rl[0], rh[0], 1u);
            k_zero_resolve<<<1, 64, 0, st>>>(lists, 1u, nodes, rl[0], rh[0]);
        }
    }
    // (offset, count) of every level's active-node list
    std::vector<std::pair<uint32_t, uint32_t>> levels_vec;
    size_t n_levels = 0;
    auto add_level = [&](uint32_t off_, uint32_t cnt_) -> bool {
        if (S.level_buf) {
            if (n_levels >= S.level_cap) return false;
            S.level_buf[n_levels][0] = off_;
            S.level_buf[n_levels][1] = cnt_;
        } else {
            levels_vec.emplace_back(off_, cnt_);
        }
        ++n_levels;
        return true;
    };
    auto level = [&](size_t i) -> std::pair<uint32_t, uint32_t> {
        return S.level_buf ? std::pair<uint32_t, uint32_t>(S.level_buf[i][0], S.level_buf[i][1]) : levels_vec[i];
    };
    uint32_t lv_off = 0, lv = 0;
    while (a > 0) {
        const uint32_t in = lv & 1u, out = in ^ 1u;
        const uint32_t* active = lists + lv_off;
        uint32_t* next = lists + lv_off + a;
        if (e == hipSuccess) e = hipMemsetAsync(&ctr->next, 0, sizeof(uint32_t), st);
        const uint32_t words = a * 3u * B * 7u;
        const uint32_t g = grid_for(nch, 1, 16384);
        k_bins_clear<<<grid_for(words, 256, 4096), 256, 0, st>>>(bins, words);
        k_bins<<<g, kWG, 0, st>>>(chunks, nch, active, nodes, rc[in], rl[in], rh[in], B, bins);
        k_split<<<grid_for(a, 64, 1u << 20), 64, 0, st>>>(a, active, nodes, bins, B, single, info, next, small, ctr, out);
        k_part_count<<<g, kWG, 0, st>>>(chunks, nch, info, nodes, rc[in], B, cnt);
        k_scan<<<1, 1024, 0, st>>>(cnt, nch, off);
        k_part<false><<<g, kWG, 0, st>>>(chunks, nch, info, nodes, off, slot_ch, B, rc[in], rl[in], rh[in], rc[out], rl[out],
                                         rh[out], tfr, tbl, d_idx, ctr);
        k_part<true><<<g, kWG, 0, st>>>(chunks, nch, info, nodes, off, slot_ch, B, rc[in], rl[in], rh[in], rc[out], rl[out],
                                        rh[out], tfr, tbl, d_idx, ctr);
        if (!read_ctr()) return WF_ERROR_DEVICE;
        const uint32_t a_next = hc.next;
        if (negzero && a_next) {
            k_zero_scan<<<g, kWG, 0, st>>>(chunks, nch, info, nodes, rl[out], rh[out], 0u);
            k_zero_resolve<<<grid_for(a_next, 64, 1u << 20), 64, 0, st>>>(next, a_next, nodes, rl[out], rh[out]);
        }
        if (!add_level(lv_off, a)) { err = "tree too deep for the fixed scratch"; return WF_ERROR_CAPACITY; }
        lv_off += a;
        a = a_next;
        ++lv;
        if (a && !make_chunks(lists + lv_off, a)) return WF_ERROR_DEVICE;
    }
    if (!read_ctr()) return WF_ERROR_DEVICE;
    const uint32_t ns = hc.small;
    if (ns)
        k_wave_build<<<grid_for(ns, kWG / 64, 65535), kWG, 0, st>>>(small, ns, nodes, rc[0], rl[0], rh[0], rc[1], rl[1], rh[1], B,
                                                                   single, blocks, d_idx, ctr);
    for (size_t i = n_levels; i-- > 0;) {
        const std::pair<uint32_t, uint32_t> l = level(i);
        k_desc<<<grid_for(l.second, 256, 1u << 20), 256, 0, st>>>(lists + l.first, l.second, nodes);
    }
    for (size_t i = 0; i < n_levels; ++i) {
        const std::pair<uint32_t, uint32_t> l = level(i);
        k_assign<<<grid_for(l.second, 256, 1u << 20), 256, 0, st>>>(lists + l.first, l.second, nodes);
    }
    if (lv_off) k_emit_upper<<<grid_for(lv_off, 256, 1u << 20), 256, 0, st>>>(lists, lv_off, nodes, d_nodes, 2u * n, ctr);
    if (ns) k_emit_small<<<grid_for(ns, kWG / 64, 65535), kWG, 0, st>>>(small, ns, nodes, blocks, d_nodes, 2u * n, ctr);
    if (e == hipSuccess) e = hipMemsetAsync(d_nodes + 1, 0, sizeof(wf_bvh_node), st);
    BNode root{};
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&root, nodes, sizeof(BNode), hipMemcpyDeviceToHost, st);
    if (!read_ctr()) return WF_ERROR_DEVICE;
    if (hc.bad_index) { err = "device build: internal index check failed"; return WF_ERROR_DEVICE; }
    res.node_count = 2u + root.desc;
    res.max_leaf = hc.max_leaf;
    for (int k = 0; k < 3; ++k) { res.bmin[k] = root.mn[k]; res.bmax[k] = root.mx[k]; }
    res.levels = lv;
    res.subtrees = ns;
    return WF_OK;
}

} // namespace wf

extern "C" WF_API wf_status wf_bvh_build_device(const float* d_tri_positions, uint32_t n, const wf_build_options* options,
                                                wf_bvh_node* d_nodes, uint64_t node_capacity, uint32_t* d_indices,
                                                uint32_t* out_node_count, void* stream)
{
    if (!d_tri_positions || !d_nodes || !d_indices || !out_node_count) return WF_ERROR_INVALID_ARGUMENT;
    if (node_capacity < 2ull * n) return WF_ERROR_CAPACITY;
    const wf_build_options opt = options ? *options : wf_build_options{8, 0};
    wf::DevBuildScratch S;
    wf::DevBuildResult res;
    std::string err;
    const wf_status s = wf::device_build_blas(d_tri_positions, n, opt, d_nodes, d_indices, S, (hipStream_t)stream, res, err);
    (void)hipStreamSynchronize((hipStream_t)stream);
    S.release();
    if (s == WF_OK) *out_node_count = res.node_count;
    return s;
}

// ---- the build without allocation -----------------------------------------
// For callers that cannot allocate device memory (graph capture, serving
// loops, harnesses that forbid it): the build's scratch is carved from one
// file-scope device array of kNoAllocScratchBytes, enough for any mesh of up
// to kNoAllocMaxTris triangles at any interval count.  Every GPU the library
// runs on holds its own copy until the process ends.  Host code never passes
// the array's name to a kernel: its address on the current device comes from
// hipGetSymbolAddress, once per device.  The array keeps the last build's
// contents; the build initialises everything it reads.  Calls on one GPU
// share the array: the calls of one process are serialised by a lock, calls
// from other processes have their own copy (their own module).
namespace {
constexpr uint32_t kNoAllocMaxTris = 65536;
constexpr size_t kNoAllocScratchBytes = 32ull << 20;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
std::mutex g_noalloc_lock;
void* g_noalloc_ptr[kNoAllocMaxDevices];                  // scratch's address per device, or null
uint32_t g_noalloc_levels[kNoAllocMaxTris][2];            // per-level bookkeeping of the build (host)
}

extern "C" WF_API wf_status wf_bvh_build_device_noalloc(const float* d_tri_positions, uint32_t n, const wf_build_options* options,
                                                        wf_bvh_node* d_nodes, uint64_t node_capacity, uint32_t* d_indices,
                                                        uint32_t* out_node_count, void* stream)
{
    if (!d_tri_positions || !d_nodes || !d_indices || !out_node_count || n == 0) return WF_ERROR_INVALID_ARGUMENT;
    if (node_capacity < 2ull * n) return WF_ERROR_CAPACITY;
    if (n > kNoAllocMaxTris) return WF_ERROR_CAPACITY;
    const wf_build_options opt = options ? *options : wf_build_options{8, 0};
    if (opt.interval_count < 2 || opt.interval_count > wf::kDevBuildMaxBins) return WF_ERROR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> guard(g_noalloc_lock);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocMaxDevices) return WF_ERROR_DEVICE;
    if (!g_noalloc_ptr[dev]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return WF_ERROR_DEVICE;
        g_noalloc_ptr[dev] = ptr;
    }
    size_t sizes[wf::kDevBuildBuffers];
    wf::device_build_sizes(n, opt.interval_count, sizes);
    wf::DevBuildScratch S;
    S.fixed = true;
    S.level_buf = g_noalloc_levels;
    S.level_cap = kNoAllocMaxTris;
    wf::DevBuf* bufs[wf::kDevBuildBuffers] = {&S.rc[0], &S.rl[0], &S.rh[0], &S.rc[1], &S.rl[1], &S.rh[1], &S.tfr, &S.tbl,
                                              &S.nodes, &S.lists, &S.small, &S.blocks, &S.info, &S.bins, &S.chunks, &S.cnt,
                                              &S.off, &S.slot_ch, &S.ctr};
    size_t at = 0;
    for (int k = 0; k < wf::kDevBuildBuffers; ++k) {
        if (at + sizes[k] > kNoAllocScratchBytes) return WF_ERROR_CAPACITY;
        bufs[k]->p = static_cast<unsigned char*>(g_noalloc_ptr[dev]) + at;
        bufs[k]->bytes = sizes[k];
        at += (sizes[k] + 255) & ~size_t(255);
    }
    wf::DevBuildResult res;
    std::string err;
    const wf_status s = wf::device_build_blas(d_tri_positions, n, opt, d_nodes, d_indices, S, (hipStream_t)stream, res, err);
    (void)hipStreamSynchronize((hipStream_t)stream);
    for (wf::DevBuf* b : bufs) { b->p = nullptr; b->bytes = 0; }   // not the scratch's to free
    if (s == WF_OK) *out_node_count = res.node_count;
    return s;
}

// A reusable device builder (bvh_builder_t, bvh_builder.h:14-37): its scratch
// grows to the largest mesh built and is kept until destroy.
struct wf_bvh_builder {
    wf::DevBuildScratch scratch;
    wf::DevBuildResult last;
    std::string err;
};

extern "C" WF_API wf_status wf_bvh_builder_create(wf_bvh_builder** out)
{
    if (!out) return WF_ERROR_INVALID_ARGUMENT;
    *out = new (std::nothrow) wf_bvh_builder();
    return *out ? WF_OK : WF_ERROR_OUT_OF_MEMORY;
}

extern "C" WF_API wf_status wf_bvh_builder_destroy(wf_bvh_builder* b)
{
    if (!b) return WF_ERROR_INVALID_ARGUMENT;
    b->scratch.release();
    delete b;
    return WF_OK;
}

extern "C" WF_API wf_status wf_bvh_builder_build_device(wf_bvh_builder* b, const float* d_tri_positions, uint32_t n,
                                                        const wf_build_options* options, wf_bvh_node* d_nodes,
                                                        uint64_t node_capacity, uint32_t* d_indices, uint32_t* out_node_count,
                                                        void* stream)
{
    if (!b || !d_tri_positions || !d_nodes || !d_indices || !out_node_count) return WF_ERROR_INVALID_ARGUMENT;
    if (node_capacity < 2ull * n) return WF_ERROR_CAPACITY;
    const wf_build_options opt = options ? *options : wf_build_options{8, 0};
    b->err.clear();
    const wf_status s = wf::device_build_blas(d_tri_positions, n, opt, d_nodes, d_indices, b->scratch, (hipStream_t)stream, b->last, b->err);
    if (s == WF_OK) *out_node_count = b->last.node_count;
    return s;
}

extern "C" WF_API wf_status wf_bvh_builder_last_build(const wf_bvh_builder* b, uint32_t* levels, uint32_t* subtrees, uint32_t* max_leaf)
{
    if (!b) return WF_ERROR_INVALID_ARGUMENT;
    if (levels) *levels = b->last.levels;
    if (subtrees) *subtrees = b->last.subtrees;
    if (max_leaf) *max_leaf = b->last.max_leaf;
    return WF_OK;
}

extern "C" WF_API const char* wf_bvh_builder_last_error(const wf_bvh_builder* b)
{
    return b ? b->err.c_str() : "null builder";
}

namespace wf {
namespace {

// 4-wide records of a BLAS node array (see kQuadBytes): record 0 holds the
// root, record r >= 1 the pair (2r, 2r+1) and the pairs of its interior
// members; interior references become record numbers (left_first / 2).
__global__ void k_make_quads(const wf_bvh_node* __restrict__ nodes, uint32_t node_count, wf_bvh_node* quads)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= node_count / 2u) return;
    auto rw = [](wf_bvh_node n) {
        if (n.prim_count == 0) n.left_first >>= 1;
        return n;
    };
    wf_bvh_node* q = quads + 6ull * r;
    wf_bvh_node z;
    for (int k = 0; k < 3; ++k) { z.aabb_min[k] = 0.0f; z.aabb_max[k] = 0.0f; }
    z.left_first = 0;
    z.prim_count = 0;
    if (r == 0) {
        q[0] = rw(nodes[0]);
        for (int k = 1; k < 6; ++k) q[k] = z;
        return;
    }
    const wf_bvh_node a = nodes[2u * r], b = nodes[2u * r + 1u];
    q[0] = rw(a);
    q[1] = rw(b);
    q[2] = a.prim_count == 0 ? rw(nodes[a.left_first]) : z;
    q[3] = a.prim_count == 0 ? rw(nodes[a.left_first + 1u]) : z;
    q[4] = b.prim_count == 0 ? rw(nodes[b.left_first]) : z;
    q[5] = b.prim_count == 0 ? rw(nodes[b.left_first + 1u]) : z;
}

// The traversal's form of a BLAS node array, in place: every leaf's
// left_first word becomes its reference word (0x80000000 | count << 27 |
// first, or 0x80000000 | first with the count in tris[first].pad0 for a leaf
// of 16 or more triangles); interior nodes keep left_first, which is already
// their reference.  prim_count is unchanged, so the reference layout stays
// recoverable (first = left_first & 0x07FFFFFF).
__global__ void k_encode_blas_refs(wf_bvh_node* nodes, uint32_t node_count, GpuTri* tris, uint32_t tri_count,
                                   uint32_t* bad)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= node_count) return;
    const uint32_t c = nodes[i].prim_count;
    if (c == 0u) return;
    const uint32_t first = nodes[i].left_first;
    if (first >= kMaxLeafRef || first >= tri_count || c > tri_count - first) { atomicOr(bad, 1u); return; }
    nodes[i].left_first = encode_ref(first, c);
    if (c >= 16u) tris[first].pad0 = c;
}

} // namespace

hipError_t launch_encode_blas_refs(wf_bvh_node* d_nodes, uint32_t node_count, GpuTri* d_tris, uint32_t tri_count,
                                   uint32_t* d_bad, hipStream_t stream)
{
    if (node_count == 0) return hipErrorInvalidValue;
    k_encode_blas_refs<<<(node_count + 255) / 256, 256, 0, stream>>>(d_nodes, node_count, d_tris, tri_count, d_bad);
    return hipGetLastError();
}

hipError_t launch_make_quads(const wf_bvh_node* d_nodes, uint32_t node_count, void* d_quads, hipStream_t stream)
{
    const uint32_t n = node_count / 2u;
    if (n == 0) return hipErrorInvalidValue;
    k_make_quads<<<(n + 255) / 256, 256, 0, stream>>>(d_nodes, node_count, (wf_bvh_node*)d_quads);
    return hipGetLastError();
}

} // namespace wf
