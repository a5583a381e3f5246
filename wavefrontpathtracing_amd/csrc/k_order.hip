// Cost-ordered queue for a launch of caller rays: the rays predicted to
// traverse longest are claimed first.
//
// A persistent extend launch claims its queue in order.  Its running time is
// the larger of the throughput-bound part and the latency of its longest
// traversal chains; a ray claimed near the end of the queue with a thousand
// dependent node visits finishes up to a millisecond after every other lane
// has gone idle.  Claiming the long rays first overlaps their chains with the
// rest of the queue (longest-processing-time-first list scheduling).  The
// order changes nothing in the results: every ray is traced exactly once and
// its hit is written at its own index.
//
// Predictor: the ray's segment inside the scene's root box (world units),
// divided by 1 + 16 d^2 / D^2, where d is the distance from the box centre to
// the ray's line and D the box diagonal (instances overlap most near the
// centre of a scene's bounds; a long chord through the middle crosses the
// most instance boxes).  The value, relative to D, is cut into 16 buckets;
// the queue holds bucket 0 (longest) first, and keeps the callers' order
// within a bucket (a stable counting sort), so coherent ray sets stay
// coherent inside each bucket.
//
// Three launches: the buckets and a per-block histogram, one scan of the
// histograms over (bucket, block), then the scatter of ray indices.  Each
// block owns a chunk of kOrderChunk consecutive rays; in the scatter each
// thread owns kOrderPerThread consecutive ones, so the order is stable
// without atomics.
#include "wf_internal.h"

#include <cfloat>

namespace wf {
namespace {

constexpr uint32_t kOrderThreads = 256;
constexpr uint32_t kOrderPerThread = 16;
constexpr uint32_t kOrderChunk = kOrderThreads * kOrderPerThread;   // 4096 rays per block

struct OrderBox {
    float mn[3], mx[3];
    float c[3];
    float inv_diag;     // 1 / |mx - mn| (0 for an empty or degenerate box)
    float inv_diag2;    // 1 / |mx - mn|^2
};

// bucket 0 = the longest predicted traversal, 15 = the shortest (and rays
// that miss the box, or whose predictor is not a number)
__device__ __forceinline__ uint32_t cost_bucket(const wf_ray& r, const OrderBox& b)
{
    const float dx = r.direction[0], dy = r.direction[1], dz = r.direction[2];
    const float len = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
    float t0 = fmaxf(r.tmin, 0.0f), t1 = r.tmax;
    const float o[3] = {r.origin[0], r.origin[1], r.origin[2]};
    const float d[3] = {dx, dy, dz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float inv = __builtin_amdgcn_rcpf(d[k]);   // a predictor: the estimate suffices
        const float ta = (b.mn[k] - o[k]) * inv, tb = (b.mx[k] - o[k]) * inv;
        t0 = fmaxf(t0, fminf(ta, tb));
        t1 = fminf(t1, fmaxf(ta, tb));
    }
    const float seg = (t1 - t0) * len;
    const float wx = o[0] - b.c[0], wy = o[1] - b.c[1], wz = o[2] - b.c[2];
    const float il = len > 0.0f ? __builtin_amdgcn_rcpf(len) : 0.0f;
    const float proj = (wx * dx + wy * dy + wz * dz) * il;
    const float dist2 = fmaxf(wx * wx + wy * wy + wz * wz - proj * proj, 0.0f);
    const float x = 16.0f * seg * b.inv_diag / (1.0f + 16.0f * dist2 * b.inv_diag2);
    // x in [0, 16): the bucket from the top; NaN and non-positive segments last
    const uint32_t q = !(x > 0.0f) ? 0u : (x >= 15.0f ? 15u : (uint32_t)x);
    return 15u - q;
}

// Pass 1: bucket[i] for every ray (one byte; rays read coalesced: ray
// blk * kOrderChunk + k * kOrderThreads + t in round k) and the chunk's
// histogram hist[b * nblocks + blk].
__global__ __launch_bounds__(kOrderThreads) void k_order_hist(const wf_ray* rays, uint32_t n, OrderBox box,
                                                              uint8_t* bucket, uint32_t* hist, uint32_t nblocks)
{
    __shared__ uint32_t s_h[kOrderBuckets];
    const uint32_t t = threadIdx.x;
    if (t < kOrderBuckets) s_h[t] = 0u;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kOrderChunk;
    for (uint32_t k = 0; k < kOrderPerThread; ++k) {
        const uint64_t i = base + (uint64_t)k * kOrderThreads + t;
        if (i < n) {
            const uint32_t b = cost_bucket(rays[i], box);
            bucket[i] = (uint8_t)b;
            atomicAdd(&s_h[b], 1u);
        }
    }
    __syncthreads();
    if (t < kOrderBuckets) hist[(size_t)t * nblocks + blockIdx.x] = s_h[t];
}

// In place: hist becomes the exclusive prefix sum over (bucket, block) in
// bucket-major order, i.e. the first queue position of each block's rays of
// each bucket.  One block of 1024 threads; thread t scans a contiguous range
// of the words, the range sums are scanned with wave shuffles.  Up to
// kScanLds words (launches of up to 2^21 rays) are staged through LDS with
// 16-byte loads and stores (total is a multiple of 16: 16 buckets x blocks).
constexpr uint32_t kScanLds = 8192;

__global__ __launch_bounds__(1024) void k_order_scan(uint32_t* hist, uint32_t total)
{
    __shared__ uint32_t s_wave[16];
    __shared__ uint4 s_h4[kScanLds / 4];
    uint32_t* s_h = reinterpret_cast<uint32_t*>(s_h4);
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    const bool staged = total <= kScanLds;
    uint32_t* h = staged ? s_h : hist;
    if (staged) {
        for (uint32_t j = t; j < total / 4u; j += 1024u) s_h4[j] = reinterpret_cast<const uint4*>(hist)[j];
        __syncthreads();
    }
    const uint32_t per = (total + 1023u) / 1024u;
    const uint32_t a = t * per < total ? t * per : total;
    const uint32_t e = a + per < total ? a + per : total;
    uint32_t sum = 0;
    for (uint32_t j = a; j < e; ++j) sum += h[j];
    uint32_t inc = sum;
#pragma unroll
    for (uint32_t off = 1; off < 64u; off <<= 1) {
        const uint32_t u = __shfl_up(inc, off, 64);
        if (lane >= off) inc += u;
    }
    if (lane == 63u) s_wave[wave] = inc;
    __syncthreads();
    uint32_t run = inc - sum;
    for (uint32_t w = 0; w < wave; ++w) run += s_wave[w];
    for (uint32_t j = a; j < e; ++j) {
        const uint32_t v = h[j];
        h[j] = run;
        run += v;
    }
    if (staged) {
        __syncthreads();
        for (uint32_t j = t; j < total / 4u; j += 1024u) reinterpret_cast<uint4*>(hist)[j] = s_h4[j];
    }
}

// Pass 2: order[pos] = ray index, stable within each bucket.  Thread t of a
// block owns the chunk's rays t * kOrderPerThread .. + kOrderPerThread - 1
// (one 16-byte load of their buckets), so the callers' order within a
// bucket is the order of (thread, ray).
__global__ __launch_bounds__(kOrderThreads) void k_order_scatter(const uint8_t* bucket, uint32_t n, const uint32_t* base,
                                                                 uint32_t nblocks, uint32_t* order)
{
    // s_c[b][t]: this thread's rays in bucket b, then their exclusive prefix
    // over the block's threads (thread t's rays follow threads 0..t-1's)
    __shared__ uint16_t s_c[kOrderBuckets][kOrderThreads];
    __shared__ uint32_t s_wave[kOrderBuckets][kOrderThreads / 64];
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    const uint64_t first = (uint64_t)blockIdx.x * kOrderChunk + (uint64_t)t * kOrderPerThread;
    uint8_t bk[kOrderPerThread];
    if (first + kOrderPerThread <= n) {
        const uint4 v = *reinterpret_cast<const uint4*>(bucket + first);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (uint32_t k = 0; k < kOrderPerThread; ++k) bk[k] = (uint8_t)(w[k / 4u] >> (8u * (k % 4u)));
    } else {
#pragma unroll
        for (uint32_t k = 0; k < kOrderPerThread; ++k) bk[k] = first + k < n ? bucket[first + k] : (uint8_t)0xFF;
    }
#pragma unroll
    for (uint32_t b = 0; b < kOrderBuckets; ++b) s_c[b][t] = 0;
#pragma unroll
    for (uint32_t k = 0; k < kOrderPerThread; ++k)
        if (bk[k] < kOrderBuckets) s_c[bk[k]][t] = (uint16_t)(s_c[bk[k]][t] + 1u);
    // exclusive scan of each bucket's counts over the block's threads
    for (uint32_t b = 0; b < kOrderBuckets; ++b) {
        const uint32_t v = s_c[b][t];
        uint32_t inc = v;
#pragma unroll
        for (uint32_t off = 1; off < 64u; off <<= 1) {
            const uint32_t u = __shfl_up(inc, off, 64);
            if (lane >= off) inc += u;
        }
        if (lane == 63u) s_wave[b][wave] = inc;
        s_c[b][t] = (uint16_t)(inc - v);
    }
    __syncthreads();
    for (uint32_t b = 0; b < kOrderBuckets; ++b) {
        uint32_t before = 0;
        for (uint32_t w = 0; w < wave; ++w) before += s_wave[b][w];
        s_c[b][t] = (uint16_t)(s_c[b][t] + before);
    }
    // positions (no other thread reads this thread's column from here on)
#pragma unroll
    for (uint32_t k = 0; k < kOrderPerThread; ++k) {
        const uint32_t b = bk[k];
        if (b < kOrderBuckets) {
            const uint32_t pos = base[(size_t)b * nblocks + blockIdx.x] + s_c[b][t];
            s_c[b][t] = (uint16_t)(s_c[b][t] + 1u);
            if (pos < n) order[pos] = (uint32_t)(first + k);   // always: the positions are a permutation of [0, n)
        }
    }
}

// bucket[i] = the cost bucket of ray i (the library's self-check)
__global__ void k_order_buckets(const wf_ray* rays, uint32_t n, OrderBox box, uint32_t* bucket)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) bucket[i] = cost_bucket(rays[i], box);
}

OrderBox make_box(const float bmin[3], const float bmax[3])
{
    OrderBox b{};
    float d2 = 0.0f;
    for (int k = 0; k < 3; ++k) {
        b.mn[k] = bmin[k];
        b.mx[k] = bmax[k];
        b.c[k] = 0.5f * (bmin[k] + bmax[k]);
        const float e = bmax[k] - bmin[k];
        d2 += e * e;
    }
    b.inv_diag = d2 > 0.0f && d2 < FLT_MAX ? 1.0f / sqrtf(d2) : 0.0f;
    b.inv_diag2 = b.inv_diag * b.inv_diag;
    return b;
}

// chunks of n rays, rounded up in 64 bits (n + kOrderChunk - 1 wraps in 32
// bits for n within a chunk of 2^32)
uint32_t order_blocks(uint32_t n) { return (uint32_t)(((uint64_t)n + kOrderChunk - 1u) / kOrderChunk); }

} // namespace

uint32_t order_scratch_words(uint32_t n)
{
    // the bucket bytes (whole chunks) and the histogram
    const uint32_t nblocks = order_blocks(n);
    return nblocks * (kOrderChunk / 4u) + kOrderBuckets * nblocks;
}

hipError_t launch_cost_order(const wf_ray* rays, uint32_t n, const float bmin[3], const float bmax[3], uint32_t* scratch,
                             uint32_t* order, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    const OrderBox b = make_box(bmin, bmax);
    const uint32_t nblocks = order_blocks(n);
    uint8_t* bucket = reinterpret_cast<uint8_t*>(scratch);
    uint32_t* hist = scratch + nblocks * (kOrderChunk / 4u);
    k_order_hist<<<nblocks, kOrderThreads, 0, stream>>>(rays, n, b, bucket, hist, nblocks);
    k_order_scan<<<1, 1024, 0, stream>>>(hist, kOrderBuckets * nblocks);
    k_order_scatter<<<nblocks, kOrderThreads, 0, stream>>>(bucket, n, hist, nblocks, order);
    return hipGetLastError();
}

hipError_t launch_cost_buckets(const wf_ray* rays, uint32_t n, const float bmin[3], const float bmax[3], uint32_t* bucket,
                               hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    k_order_buckets<<<(uint32_t)(((uint64_t)n + 255u) / 256u), 256, 0, stream>>>(rays, n, make_box(bmin, bmax), bucket);
    return hipGetLastError();
}

} // namespace wf

// Self-check of the cost-ordered queue (wf_selftest_cost_order, wf.h): the
// order and each ray's bucket for device rays.
extern "C" __attribute__((visibility("default"))) wf_status wf_selftest_cost_order(const wf_ray* d_rays, uint32_t n,
                                                                                    const float bmin[3], const float bmax[3],
                                                                                    uint32_t* d_order, uint32_t* d_bucket)
{
    if (n && (!d_rays || !d_order || !d_bucket)) return WF_ERROR_INVALID_ARGUMENT;
    if (!bmin || !bmax) return WF_ERROR_INVALID_ARGUMENT;
    if (n == 0) return WF_OK;
    void* scratch = nullptr;
    const size_t bytes = 4ull * wf::order_scratch_words(n);
    if (hipMalloc(&scratch, bytes) != hipSuccess) return WF_ERROR_OUT_OF_MEMORY;
    hipError_t e = wf::launch_cost_order(d_rays, n, bmin, bmax, (uint32_t*)scratch, d_order, nullptr);
    if (e == hipSuccess) e = wf::launch_cost_buckets(d_rays, n, bmin, bmax, d_bucket, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    hipFree(scratch);
    return e == hipSuccess ? WF_OK : WF_ERROR_DEVICE;
}
