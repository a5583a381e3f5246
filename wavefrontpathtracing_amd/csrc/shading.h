// Shading of one path vertex and the helpers the stage kernels share: the
// build's RNG (PCG32 keyed on path and bounce), cosine sampling
// (sample.hlsl:13-27,45-54), bilinear texture sampling, glm's vector
// arithmetic order, and the wave64 queue appends.  Used by the wavefront
// shade kernel (forward and adjoint) and by the megakernel, so both render
// the same paths with the same operation sequence.
#pragma once

#include "stages.h"

#include <cfloat>

#pragma clang fp contract(off)

namespace wf {
namespace shading {

constexpr float kInvPi = 0.31830988f;   // common.hlsl:23
constexpr float kInvTwoPi = 0.15915494f;   // common.hlsl:24
constexpr float kRayMinT = 1e-8f;

// ---- RNG: PCG32 (O'Neill 2014) seeded per (path, bounce) ------------------
struct Pcg { uint64_t state, inc; };
__device__ __forceinline__ uint32_t pcg_next(Pcg& r)
{
    const uint64_t old = r.state;
    r.state = old * 6364136223846793005ull + r.inc;
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}
__device__ __forceinline__ void path_rng(uint32_t path, uint32_t bounce, uint32_t seed, float u[6])
{
    Pcg r;
    r.state = 0u;
    r.inc = ((uint64_t)bounce << 1u) | 1u;
    pcg_next(r);
    r.state += ((uint64_t)seed << 32) | path;
    pcg_next(r);
#pragma unroll
    for (int k = 0; k < 6; ++k) u[k] = (float)(pcg_next(r) >> 8) * (1.0f / 16777216.0f);
}

// ---- sin/cos of 2*pi*u (quadrant + Taylor, the same operation sequence as
// the host restatement) ----------------------------------------------------------
__device__ __forceinline__ void sincos_turns(float u, float& s, float& c)
{
    const float q4 = u * 4.0f;
    const float qf = floorf(q4);
    const float f = q4 - qf;
    const int q = ((int)qf) & 3;
    const float x = f * 1.57079632679489662f;
    const float y = x * x;
    float p = (float)(1.0 / 6227020800.0);
    p = p * y - (float)(1.0 / 39916800.0);
    p = p * y + (float)(1.0 / 362880.0);
    p = p * y - (float)(1.0 / 5040.0);
    p = p * y + (float)(1.0 / 120.0);
    p = p * y - (float)(1.0 / 6.0);
    p = p * y + 1.0f;
    const float sx = x * p;
    float r = -(float)(1.0 / 87178291200.0);
    r = r * y + (float)(1.0 / 479001600.0);
    r = r * y - (float)(1.0 / 3628800.0);
    r = r * y + (float)(1.0 / 40320.0);
    r = r * y - (float)(1.0 / 720.0);
    r = r * y + (float)(1.0 / 24.0);
    r = r * y - 0.5f;
    r = r * y + 1.0f;
    const float cx = r;
    switch (q) {
    case 0: s = sx; c = cx; break;
    case 1: s = cx; c = -sx; break;
    case 2: s = -sx; c = -cx; break;
    default: s = -cx; c = sx; break;
    }
}

struct V3 { float x, y, z; };
__device__ __forceinline__ float dot3(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 normalize3(V3 v)
{
    const float s = 1.0f / sqrtf(dot3(v, v));
    return V3{v.x * s, v.y * s, v.z * s};
}
__device__ __forceinline__ V3 cross3(V3 a, V3 b)
{
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// glm mat4 * vec4 association: (c0*x + c1*y) + (c2*z + c3*w), xyz only
__device__ __forceinline__ V3 xform(const float* m, V3 v, float w)
{
    V3 o;
    o.x = (m[0] * v.x + m[4] * v.y) + (m[8] * v.z + m[12] * w);
    o.y = (m[1] * v.x + m[5] * v.y) + (m[9] * v.z + m[13] * w);
    o.z = (m[2] * v.x + m[6] * v.y) + (m[10] * v.z + m[14] * w);
    return o;
}
__device__ __forceinline__ float interp(float a, float b, float c, float bx, float by) { return (a + bx * (b - a)) + by * (c - a); }

// cosine_weighted_hemisphere_sample + create_orthonormal_basis (sample.hlsl:13-27,45-54)
__device__ __forceinline__ V3 cosine_sample(V3 n, float r0, float r1)
{
    const float sin_theta = sqrtf(r0);
    float sp, cp;
    sincos_turns(r1, sp, cp);
    const V3 v = normalize3(V3{sin_theta * cp, sqrtf(1.0f - r0), sin_theta * sp});
    const V3 t0 = (fabsf(n.x) > fabsf(n.z)) ? V3{-n.y, n.x, 0.0f} : V3{0.0f, -n.z, n.y};
    const V3 t = normalize3(t0);
    const V3 b = cross3(n, t);
    return V3{(v.x * t.x + v.y * n.x) + v.z * b.x, (v.x * t.y + v.y * n.y) + v.z * b.y, (v.x * t.z + v.y * n.z) + v.z * b.z};
}

// uniform_hemisphere_sample + create_orthonormal_basis (sample.hlsl:13-43)
__device__ __forceinline__ V3 uniform_sample(V3 n, float r0, float r1)
{
    const float sin_theta = sqrtf(1.0f - r0 * r0);
    float sp, cp;
    sincos_turns(r1, sp, cp);
    const V3 v = normalize3(V3{sin_theta * cp, r0, sin_theta * sp});
    const V3 t0 = (fabsf(n.x) > fabsf(n.z)) ? V3{-n.y, n.x, 0.0f} : V3{0.0f, -n.z, n.y};
    const V3 t = normalize3(t0);
    const V3 b = cross3(n, t);
    return V3{(v.x * t.x + v.y * n.x) + v.z * b.x, (v.x * t.y + v.y * n.y) + v.z * b.y, (v.x * t.z + v.y * n.z) + v.z * b.z};
}

__device__ __forceinline__ V3 tex_sample(const GpuTexture& tx, float u, float v)
{
    float fx = u * (float)tx.width - 0.5f, fy = v * (float)tx.height - 0.5f;
    fx = fminf(fmaxf(fx, -1e9f), 1e9f);
    fy = fminf(fmaxf(fy, -1e9f), 1e9f);
    const float fx0 = floorf(fx), fy0 = floorf(fy);
    const float ax = fx - fx0, ay = fy - fy0;
    const int64_t ix = (int64_t)fx0, iy = (int64_t)fy0;
    const int64_t w = tx.width, h = tx.height;
    const uint32_t x0 = (uint32_t)(((ix % w) + w) % w), y0 = (uint32_t)(((iy % h) + h) % h);
    const uint32_t x1 = (x0 + 1u == (uint32_t)w) ? 0u : x0 + 1u, y1 = (y0 + 1u == (uint32_t)h) ? 0u : y0 + 1u;
    const float4* t = reinterpret_cast<const float4*>(tx.texels);
    const float4 a = t[(size_t)y0 * w + x0], b = t[(size_t)y0 * w + x1];
    const float4 c = t[(size_t)y1 * w + x0], d = t[(size_t)y1 * w + x1];
    V3 o;
    float top, bot;
    top = a.x + ax * (b.x - a.x); bot = c.x + ax * (d.x - c.x); o.x = top + ay * (bot - top);
    top = a.y + ax * (b.y - a.y); bot = c.y + ax * (d.y - c.y); o.y = top + ay * (bot - top);
    top = a.z + ax * (b.z - a.z); bot = c.z + ax * (d.z - c.z); o.z = top + ay * (bot - top);
    return o;
}

// one atomic per wave64 for all lanes appending to the same queue
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred)
{
    const uint64_t mask = __ballot(pred);
    uint32_t slot = 0;
    if (pred) {
        const uint32_t lane = __lane_id();
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
        const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
        base = __shfl(base, (int)leader, 64);
        slot = base + rank;
    }
    return slot;
}

// atomicAdd(&arr[key], v) for every lane with `pred`, combined per distinct
// key: the wave sums the values of the lanes that share the first pending
// lane's key and issues one atomic for them (the emission gradients of a
// frame all land on a handful of emitter instances).  Called by all lanes.
__device__ __forceinline__ void wave_add_keyed(float* arr, uint32_t key, float v, bool pred)
{
    const uint32_t lane = __lane_id();
    uint64_t pending = __ballot(pred);
    while (pending) {
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)pending) - 1u;
        const uint32_t k = __shfl(key, (int)leader, 64);
        const bool mine = pred && key == k;
        const uint64_t sel = __ballot(mine);
        float sum = mine ? v : 0.0f;
        for (int off = 32; off > 0; off >>= 1) sum = sum + __shfl_xor(sum, off, 64);
        if (lane == leader) atomicAdd(arr + k, sum);
        pending &= ~sel;
        pred = pred && !mine;
    }
}

__device__ __forceinline__ void store_ray(wf_ray* q, uint32_t slot, V3 o, V3 d, float tmin, float tmax)
{
    float4* p = reinterpret_cast<float4*>(q + slot);
    p[0] = make_float4(o.x, o.y, o.z, tmin);
    p[1] = make_float4(d.x, d.y, d.z, tmax);
}

// scene slot holding global path id p (slots are sorted by path_off)
__device__ __forceinline__ uint32_t scene_of_path(const FrameArgs& A, uint32_t p)
{
    uint32_t lo = A.first_scene, hi = A.first_scene + A.n_scenes - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (A.scenes[mid].path_off <= p) lo = mid; else hi = mid - 1u;
    }
    return lo;
}

// The order in which camera rays enter the first queue: 8x8 pixel tiles
// (one wave64 each) over the part of the image the tiles cover, then the
// right strip and the bottom strip row by row.  A bijection on [0, W*H).
// Only the queue order changes: a path keeps its id (and so its random
// numbers and its pixel), so the image does not depend on it.
__device__ __forceinline__ uint32_t tile_pixel(uint32_t k, uint32_t W, uint32_t H)
{
    const uint32_t W8 = W & ~7u, H8 = H & ~7u;
    const uint32_t n_full = W8 * H8;
    if (k < n_full) {
        const uint32_t tile = k >> 6, local = k & 63u;
        const uint32_t tiles_per_row = W8 >> 3;
        return ((tile / tiles_per_row) * 8u + (local >> 3)) * W + (tile % tiles_per_row) * 8u + (local & 7u);
    }
    k -= n_full;
    const uint32_t rw = W - W8;
    const uint32_t n_right = rw * H8;
    if (k < n_right) return (k / rw) * W + W8 + k % rw;
    return H8 * W + (k - n_right);
}

// the path of first-queue position i (0-based in the frame) and its scene
__device__ __forceinline__ uint32_t path_of_position(const FrameArgs& A, uint32_t i, uint32_t& scene)
{
    const uint32_t lin = A.path_base + i;
    scene = scene_of_path(A, lin);
    const GpuScene& sc = A.scenes[scene];
    const uint32_t q = lin - sc.path_off;
    const uint32_t sample = q / sc.n_pix, k = q - sample * sc.n_pix;
    return sc.path_off + sample * sc.n_pix + tile_pixel(k, sc.width, sc.height);
}

__device__ __forceinline__ uint32_t scene_of_pixel(const FrameArgs& A, uint32_t q)
{
    uint32_t lo = A.first_scene, hi = A.first_scene + A.n_scenes - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (A.scenes[mid].pix_off <= q) lo = mid; else hi = mid - 1u;
    }
    return lo;
}

// bilinear corners and fractions of tex_sample, for scattering gradients
__device__ __forceinline__ void tex_corners(const GpuTexture& tx, float u, float v, uint32_t c[4], float& ax, float& ay)
{
    float fx = u * (float)tx.width - 0.5f, fy = v * (float)tx.height - 0.5f;
    fx = fminf(fmaxf(fx, -1e9f), 1e9f);
    fy = fminf(fmaxf(fy, -1e9f), 1e9f);
    const float fx0 = floorf(fx), fy0 = floorf(fy);
    ax = fx - fx0;
    ay = fy - fy0;
    const int64_t ix = (int64_t)fx0, iy = (int64_t)fy0;
    const int64_t w = tx.width, h = tx.height;
    const uint32_t x0 = (uint32_t)(((ix % w) + w) % w), y0 = (uint32_t)(((iy % h) + h) % h);
    const uint32_t x1 = (x0 + 1u == (uint32_t)w) ? 0u : x0 + 1u, y1 = (y0 + 1u == (uint32_t)h) ? 0u : y0 + 1u;
    c[0] = y0 * (uint32_t)w + x0; c[1] = y0 * (uint32_t)w + x1;
    c[2] = y1 * (uint32_t)w + x0; c[3] = y1 * (uint32_t)w + x1;
}

// make_primary_ray + make_ray (common.hlsl:145-169): the camera ray of a
// scene's pixel, origin nudged by 1e-8 along the direction
__device__ __forceinline__ void camera_ray(const GpuScene& sc, uint32_t pixel, V3& o, V3& dw)
{
    const uint32_t px = pixel % sc.width, py = pixel / sc.width;
    const float* c2v = sc.view.clip_to_view;
    const float* v2w = sc.view.view_to_world;
    const float uvx = ((float)px + 0.5f) / sc.view.render_dim[0];
    float uvy = ((float)py + 0.5f) / sc.view.render_dim[1];
    uvy = 1.0f - uvy;
    const V3 cv = xform(c2v, V3{2.0f * uvx - 1.0f, 2.0f * uvy - 1.0f, 1.0f}, 1.0f);
    const V3 dv = normalize3(cv);
    dw = xform(v2w, dv, 0.0f);
    const V3 ow = xform(v2w, V3{0.0f, 0.0f, 0.0f}, 1.0f);
    o = V3{ow.x + dw.x * kRayMinT, ow.y + dw.y * kRayMinT, ow.z + dw.z * kRayMinT};
}

// What shading one vertex produces: the continuation ray (spawn), the
// next-event shadow ray with its contribution (shadow), and for the adjoint
// the shadow contribution per unit emission strength, its emitter, and the
// emission gradient of an emitter hit (em_*).
struct VertexOut {
    bool spawn = false, shadow = false;
    V3 new_o{0, 0, 0}, new_d{0, 0, 0}, sh_o{0, 0, 0}, sh_d{0, 0, 0};
    float sh_tmax = 0.0f;
    float4 contrib = make_float4(0, 0, 0, 0), unit = make_float4(0, 0, 0, 0);
    uint32_t light = 0, em_key = 0;
    float em_g = 0.0f;
    bool em_has = false;
};

// shade.hlsl:42-174 with the build's RNG, Russian roulette and next-event
// estimation (the host restatement has the same operation order), for path
// p of scene `scene` at bounce d with closest hit `hit`.  thr is the path
// throughput; e4 the path energy (forward) or the remaining estimate lrem
// (ADJ).
//
// ADJ = true is the path-replay adjoint (Vicini, Speierer, Jakob 2021): the
// same path is replayed with the same random numbers; instead of adding
// contributions to the path's energy it subtracts them from lrem, and since
// every contribution is linear in the albedo of each vertex it passes,
// dEstimate/dalbedo_d = lrem / albedo_d at vertex d, scattered to the
// texture's bilinear corners (dlt = dLoss/dEstimate); emission gradients are
// the collected contributions per unit emission strength.
// Texel of an equirect environment in direction d (sample_hdr_env,
// shade.hlsl:36-40, with direction_to_equirect_uv, sample.hlsl:4-11):
// uv = (atan2(z, x), asin(-y)) * (0.1591, 0.3183) + 0.5, a point load at
// uint2(uv * size).  0.1591 < 1/(2 pi) and 0.3183 < 1/pi keep uv below 1 for
// every finite direction; other inputs are clamped into the texture.
__device__ __forceinline__ uint32_t env_texel(const GpuTexture& T, const V3& d)
{
    const float u = atan2f(d.z, d.x) * 0.1591f + 0.5f;
    const float v = asinf(-d.y) * 0.3183f + 0.5f;
    const float fx = u * (float)T.width, fy = v * (float)T.height;
    uint32_t x = fx >= 1.0f ? (uint32_t)fx : 0u, y = fy >= 1.0f ? (uint32_t)fy : 0u;
    x = x < T.width ? x : T.width - 1u;
    y = y < T.height ? y : T.height - 1u;
    return y * T.width + x;
}

// murmur_mix (common.hlsl:276-284): MurmurHash3's 32-bit finaliser
__device__ __forceinline__ uint32_t murmur_mix(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

// int_to_color (common.hlsl:286-295)
__device__ __forceinline__ V3 int_to_color(uint32_t v)
{
    const uint32_t h = murmur_mix(v);
    const float s = 1.0f / 255.0f;
    return V3{(float)(h & 255u) * s, (float)((h >> 8) & 255u) * s, (float)((h >> 16) & 255u) * s};
}

// The debug view modes (shade.hlsl:140-153): one property of the camera
// ray's hit replaces its radiance.  A miss sees what the reference reads for
// an invalid instance (zeroed records): its ids, barycentrics (0, 0) and t;
// zero geometry and material values; a world normal of normalize(0) (NaN,
// which the post-process paints magenta).  The material normal and
// metallic-roughness views of a miss, which the reference takes from
// whatever texture sits in descriptor slot 0, are 0 here.
__device__ __forceinline__ V3 view_color(const FrameArgs& A, uint32_t scene, const wf_hit& hit)
{
    const uint32_t m = A.view_mode;
    if (m == WF_RENDER_VIEW_MODE_GEOMETRY_INSTANCE) return int_to_color(hit.instance_idx);
    if (m == WF_RENDER_VIEW_MODE_GEOMETRY_PRIMITIVE) return int_to_color(hit.primitive_idx);
    if (m == WF_RENDER_VIEW_MODE_GEOMETRY_BARYCENTRICS) return V3{hit.bary[0], hit.bary[1], 0.0f};
    if (m == WF_RENDER_VIEW_MODE_RENDER_TARGET_DEPTH) {
        const float z = hit.t / A.scenes[scene].view.far_plane;
        return V3{z, z, z};
    }
    if (hit.instance_idx == 0xFFFFFFFFu) {
        if (m == WF_RENDER_VIEW_MODE_WORLD_NORMAL) {
            const V3 n = normalize3(V3{0.0f, 0.0f, 0.0f});
            return V3{fabsf(n.x), fabsf(n.y), fabsf(n.z)};
        }
        return V3{0.0f, 0.0f, 0.0f};
    }
    const GpuShadeInstance* si = A.shade_inst + hit.instance_idx;
    const GpuShadeTri* tri = si->tris + hit.primitive_idx;
    const float bx = hit.bary[0], by = hit.bary[1];
    const float tu = interp(tri->uv[0][0], tri->uv[1][0], tri->uv[2][0], bx, by);
    const float tv = interp(tri->uv[0][1], tri->uv[1][1], tri->uv[2][1], bx, by);
    switch (m) {
    case WF_RENDER_VIEW_MODE_GEOMETRY_NORMAL:
    case WF_RENDER_VIEW_MODE_WORLD_NORMAL: {
        const V3 ln{interp(tri->n[0][0], tri->n[1][0], tri->n[2][0], bx, by),
                    interp(tri->n[0][1], tri->n[1][1], tri->n[2][1], bx, by),
                    interp(tri->n[0][2], tri->n[1][2], tri->n[2][2], bx, by)};
        const V3 n = m == WF_RENDER_VIEW_MODE_WORLD_NORMAL ? normalize3(xform(si->local_to_world, ln, 0.0f)) : ln;
        return V3{fabsf(n.x), fabsf(n.y), fabsf(n.z)};
    }
    case WF_RENDER_VIEW_MODE_GEOMETRY_UV: return V3{tu, tv, 0.0f};
    case WF_RENDER_VIEW_MODE_MATERIAL_BASE_COLOR: {
        V3 c{si->base_color[0], si->base_color[1], si->base_color[2]};
        if (si->base_color_tex >= 0) {
            const V3 tc = tex_sample(A.textures[si->base_color_tex], tu, tv);
            c = V3{c.x * tc.x, c.y * tc.y, c.z * tc.z};
        }
        return c;
    }
    case WF_RENDER_VIEW_MODE_MATERIAL_NORMAL: {
        // the default normal texture is RGBA8 (127, 127, 255, 255) (renderer.cpp:256-262)
        const V3 n = si->normal_tex >= 0 ? tex_sample(A.textures[si->normal_tex], tu, tv)
                                         : V3{127.0f / 255.0f, 127.0f / 255.0f, 1.0f};
        return V3{fabsf(n.x), fabsf(n.y), fabsf(n.z)};
    }
    case WF_RENDER_VIEW_MODE_MATERIAL_METALLIC_ROUGHNESS: {
        // (0, roughness = g, metallic = b); the default texture is RG8 (255, 255) (renderer.cpp:264-269)
        const V3 t = si->metal_rough_tex >= 0 ? tex_sample(A.textures[si->metal_rough_tex], tu, tv) : V3{1.0f, 1.0f, 0.0f};
        return V3{0.0f, t.y, t.z};
    }
    case WF_RENDER_VIEW_MODE_MATERIAL_EMISSIVE: {
        const float es = si->emissive_strength;
        V3 etc{1.0f, 1.0f, 1.0f};
        if (si->emissive_tex >= 0) etc = tex_sample(A.textures[si->emissive_tex], tu, tv);
        return V3{(si->emissive_factor[0] * es) * etc.x, (si->emissive_factor[1] * es) * etc.y,
                  (si->emissive_factor[2] * es) * etc.z};
    }
    default: return V3{0.0f, 0.0f, 0.0f};
    }
}

// dir: the direction of the ray that reached this vertex (read for misses only)
template <bool ADJ>
__device__ __forceinline__ void shade_vertex(const FrameArgs& A, uint32_t p, uint32_t scene, uint32_t d, uint32_t seed,
                                             const wf_hit& hit, const V3& dir, V3& thr, float4& e4, const float4& dlt,
                                             VertexOut& o)
{
    if (!ADJ && A.view_mode != 0u) {
        // a debug view replaces the camera ray's radiance and ends the path
        const V3 c = view_color(A, scene, hit);
        e4.x = c.x;
        e4.y = c.y;
        e4.z = c.z;
        return;
    }
    float u[6];
    path_rng(p, d, seed, u);
    bool done = false;
    if (hit.instance_idx == 0xFFFFFFFFu) {
        // environment radiance: throughput * hdr_env_strength * env (shade.hlsl:62-63)
        float ex, ey, ez;
        if (A.env_tex >= 0) {
            const GpuTexture& T = A.textures[A.env_tex];
            const float* c = T.texels + 4ull * env_texel(T, dir);
            ex = (thr.x * A.env_strength) * c[0];
            ey = (thr.y * A.env_strength) * c[1];
            ez = (thr.z * A.env_strength) * c[2];
        } else {
            ex = thr.x * A.env[0];
            ey = thr.y * A.env[1];
            ez = thr.z * A.env[2];
        }
        if (ADJ) {
            e4.x = e4.x - ex;
            e4.y = e4.y - ey;
            e4.z = e4.z - ez;
        } else {
            e4.x = e4.x + ex;
            e4.y = e4.y + ey;
            e4.z = e4.z + ez;
        }
        done = true;
    }
    if (!done) {
        const GpuShadeInstance* si = A.shade_inst + hit.instance_idx;
        const float* l2w = si->local_to_world;
        const GpuShadeTri* tri = si->tris + hit.primitive_idx;
        const float bx = hit.bary[0], by = hit.bary[1];
        const V3 lp{interp(tri->p[0][0], tri->p[1][0], tri->p[2][0], bx, by),
                    interp(tri->p[0][1], tri->p[1][1], tri->p[2][1], bx, by),
                    interp(tri->p[0][2], tri->p[1][2], tri->p[2][2], bx, by)};
        const V3 ln{interp(tri->n[0][0], tri->n[1][0], tri->n[2][0], bx, by),
                    interp(tri->n[0][1], tri->n[1][1], tri->n[2][1], bx, by),
                    interp(tri->n[0][2], tri->n[1][2], tri->n[2][2], bx, by)};
        const float tu = interp(tri->uv[0][0], tri->uv[1][0], tri->uv[2][0], bx, by);
        const float tv = interp(tri->uv[0][1], tri->uv[1][1], tri->uv[2][1], bx, by);
        const V3 P = xform(l2w, lp, 1.0f);
        const V3 N = normalize3(xform(l2w, ln, 0.0f));
        V3 albedo{si->base_color[0], si->base_color[1], si->base_color[2]};
        if (si->base_color_tex >= 0) {
            const V3 tc = tex_sample(A.textures[si->base_color_tex], tu, tv);
            albedo = V3{albedo.x * tc.x, albedo.y * tc.y, albedo.z * tc.z};
        }
        // emission = strength * factor * emissive texel (sample_material, material.hlsl:35-37)
        const float es = si->emissive_strength;
        V3 etc{1.0f, 1.0f, 1.0f};
        if (si->emissive_tex >= 0) etc = tex_sample(A.textures[si->emissive_tex], tu, tv);
        const V3 ef{si->emissive_factor[0] * etc.x, si->emissive_factor[1] * etc.y, si->emissive_factor[2] * etc.z};
        const V3 em{(si->emissive_factor[0] * es) * etc.x, (si->emissive_factor[1] * es) * etc.y,
                    (si->emissive_factor[2] * es) * etc.z};
        if (em.x != 0.0f || em.y != 0.0f || em.z != 0.0f) {
            if (!A.nee || d == 0) {
                if (ADJ) {
                    e4.x = e4.x - thr.x * em.x;
                    e4.y = e4.y - thr.y * em.y;
                    e4.z = e4.z - thr.z * em.z;
                    // d emission / d strength = factor * texel
                    o.em_g = (dlt.x * (thr.x * ef.x) + dlt.y * (thr.y * ef.y)) + dlt.z * (thr.z * ef.z);
                    o.em_key = hit.instance_idx;
                    o.em_has = o.em_g != 0.0f;
                } else {
                    e4.x = e4.x + thr.x * em.x;
                    e4.y = e4.y + thr.y * em.y;
                    e4.z = e4.z + thr.z * em.z;
                }
            }
            done = true;
        }
        if (ADJ && !done && si->base_color_tex == A.grad_tex_slot && si->base_color_tex >= 0) {
            // d estimate / d texel: (lrem / albedo) * base_color * bilinear weight
            uint32_t c[4];
            float ax, ay;
            tex_corners(A.textures[si->base_color_tex], tu, tv, c, ax, ay);
            const float wts[4] = {(1.0f - ax) * (1.0f - ay), ax * (1.0f - ay), (1.0f - ax) * ay, ax * ay};
            const float g3[3] = {albedo.x > 0.0f ? dlt.x * e4.x / albedo.x * si->base_color[0] : 0.0f,
                                 albedo.y > 0.0f ? dlt.y * e4.y / albedo.y * si->base_color[1] : 0.0f,
                                 albedo.z > 0.0f ? dlt.z * e4.z / albedo.z * si->base_color[2] : 0.0f};
            if (g3[0] != 0.0f || g3[1] != 0.0f || g3[2] != 0.0f) {
                for (int j = 0; j < 4; ++j) {
                    float* gt = reinterpret_cast<float*>(A.grad_tex + c[j]);
                    atomicAdd(gt + 0, g3[0] * wts[j]);
                    atomicAdd(gt + 1, g3[1] * wts[j]);
                    atomicAdd(gt + 2, g3[2] * wts[j]);
                }
            }
        }
        if (!done && A.nee && d < A.max_bounces) {
            const GpuScene& sc = A.scenes[scene];
            if (sc.em_count > 0) {
                uint32_t lo = 0, hi = sc.em_count - 1u;
                const uint4* em_tab = A.emitters + sc.em_off;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (__uint_as_float(em_tab[mid].z) > u[3]) hi = mid; else lo = mid + 1u;
                }
                const uint4 er = em_tab[lo];
                const GpuShadeInstance* li = A.shade_inst + er.x;
                const GpuShadeTri* lt = li->tris + er.y;
                const V3 W0 = xform(li->local_to_world, V3{lt->p[0][0], lt->p[0][1], lt->p[0][2]}, 1.0f);
                const V3 W1 = xform(li->local_to_world, V3{lt->p[1][0], lt->p[1][1], lt->p[1][2]}, 1.0f);
                const V3 W2 = xform(li->local_to_world, V3{lt->p[2][0], lt->p[2][1], lt->p[2][2]}, 1.0f);
                const float su = sqrtf(u[4]);
                const float bv = su * (1.0f - u[5]), bw = su * u[5];
                const V3 PL{interp(W0.x, W1.x, W2.x, bv, bw), interp(W0.y, W1.y, W2.y, bv, bw), interp(W0.z, W1.z, W2.z, bv, bw)};
                const V3 dl{PL.x - P.x, PL.y - P.y, PL.z - P.z};
                const V3 e1{W1.x - W0.x, W1.y - W0.y, W1.z - W0.z};
                const V3 e2{W2.x - W0.x, W2.y - W0.y, W2.z - W0.z};
                const float dist2 = dot3(dl, dl);
                const float dist = sqrtf(dist2);
                const float inv_d = 1.0f / dist;
                const V3 wi{dl.x * inv_d, dl.y * inv_d, dl.z * inv_d};
                const V3 nl = normalize3(cross3(e1, e2));
                const float cos_s = dot3(N, wi);
                const float cos_l = -dot3(nl, wi);
                if (cos_s > 0.0f && cos_l > 0.0f && dist2 > 0.0f) {
                    const float G = (cos_s * cos_l) / dist2;
                    const float wgt = G * sc.em_area;
                    const float ls = li->emissive_strength;
                    // the light's emissive texel at the sampled point
                    V3 ltc{1.0f, 1.0f, 1.0f};
                    if (li->emissive_tex >= 0)
                        ltc = tex_sample(A.textures[li->emissive_tex], interp(lt->uv[0][0], lt->uv[1][0], lt->uv[2][0], bv, bw),
                                         interp(lt->uv[0][1], lt->uv[1][1], lt->uv[2][1], bv, bw));
                    o.contrib.x = (((thr.x * albedo.x) * kInvPi) * ((li->emissive_factor[0] * ls) * ltc.x)) * wgt;
                    o.contrib.y = (((thr.y * albedo.y) * kInvPi) * ((li->emissive_factor[1] * ls) * ltc.y)) * wgt;
                    o.contrib.z = (((thr.z * albedo.z) * kInvPi) * ((li->emissive_factor[2] * ls) * ltc.z)) * wgt;
                    if (ADJ) {
                        o.unit.x = (((thr.x * albedo.x) * kInvPi) * (li->emissive_factor[0] * ltc.x)) * wgt;
                        o.unit.y = (((thr.y * albedo.y) * kInvPi) * (li->emissive_factor[1] * ltc.y)) * wgt;
                        o.unit.z = (((thr.z * albedo.z) * kInvPi) * (li->emissive_factor[2] * ltc.z)) * wgt;
                        o.light = er.x;
                    }
                    o.sh_o = V3{P.x + wi.x * kRayMinT, P.y + wi.y * kRayMinT, P.z + wi.z * kRayMinT};
                    o.sh_d = wi;
                    o.sh_tmax = dist * 0.9999f;
                    o.shadow = true;
                }
            }
        }
        if (!done) {
            // shade.hlsl:108-135: cosine-weighted (default) or uniform hemisphere sample
            const V3 L = A.uniform_diffuse ? uniform_sample(N, u[0], u[1]) : cosine_sample(N, u[0], u[1]);
            float NoL = dot3(N, L);
            if (NoL < 0.0f) NoL = 0.0f;
            if (!(NoL > 0.0f)) {
                done = true;
            } else {
                const float pdf = A.uniform_diffuse ? kInvTwoPi : NoL * kInvPi;
                const float inv_pdf = 1.0f / pdf;
                thr.x = thr.x * ((NoL * (albedo.x * kInvPi)) * inv_pdf);
                thr.y = thr.y * ((NoL * (albedo.y * kInvPi)) * inv_pdf);
                thr.z = thr.z * ((NoL * (albedo.z * kInvPi)) * inv_pdf);
                if (d >= A.rr_start) {
                    if (u[2] >= A.rr_survival) done = true;
                    else { thr.x = thr.x / A.rr_survival; thr.y = thr.y / A.rr_survival; thr.z = thr.z / A.rr_survival; }
                }
                if (!done && d < A.max_bounces) {
                    o.spawn = true;
                    o.new_o = V3{P.x + L.x * kRayMinT, P.y + L.y * kRayMinT, P.z + L.z * kRayMinT};
                    o.new_d = L;
                }
            }
        }
    }
}

} // namespace shading
} // namespace wf
