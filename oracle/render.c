/*
 * Oracle: the wavefront render loop, per path.  TEST INFRASTRUCTURE ONLY
 * (see oracle.h).
 *
 * Follows generate.hlsl, extend.hlsl and shade.hlsl:42-174 (get_hit_surface
 * common.hlsl:228-245, sample_material material.hlsl:15-39,
 * cosine_weighted_hemisphere_sample sample.hlsl:13-27,45-54) with this
 * build's documented changes:
 *   - random numbers: PCG32 (O'Neill 2014) seeded per (path, bounce) instead
 *     of a Wang hash of frame_seed + queue slot (shade.hlsl:80), so a path's
 *     result does not depend on the order of the compacted queue;
 *     float = (u32 >> 8) * 2^-24 in [0, 1);
 *   - cos/sin of 2*pi*u are evaluated in turns by a shared Taylor polynomial;
 *   - Russian roulette from bounce rr_start with a constant survival
 *     probability (no dependence on albedo);
 *   - optional next-event estimation (the reference's connect stage is empty,
 *     connect.hlsl:1): one shadow ray per non-emissive vertex d < max_bounces
 *     to a triangle of the emitter list chosen by area; emission found by a
 *     bounce ray (d > 0) is then not added again;
 *   - the environment is a constant radiance (the reference's fallback 1x1
 *     white texture times hdr_env_strength, renderer.cpp:605-609);
 *   - a bounce with N.L <= 0 ends the path (the reference divides by a zero
 *     pdf there).
 * Paths are independent, so the oracle traces each path to completion; the
 * per-path arithmetic is the same as the stage-by-stage GPU loop.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define INV_PI 0.31830988f     /* common.hlsl:23 */
#define INV_TWO_PI 0.15915494f /* common.hlsl:24 */
#define RAY_MIN_T 1e-8f

int orc_trace_one(const orc_scene* s, const orc_ray* r, orc_hit* h, int any);

/* ---- RNG ------------------------------------------------------------------ */
typedef struct { uint64_t state, inc; } pcg32;

static uint32_t pcg_next(pcg32* r)
{
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ull + r->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

static void pcg_seed(pcg32* r, uint64_t initstate, uint64_t initseq)
{
    r->state = 0u;
    r->inc = (initseq << 1u) | 1u;
    pcg_next(r);
    r->state += initstate;
    pcg_next(r);
}

uint32_t orc_pcg32_first(uint64_t initstate, uint64_t initseq)
{
    pcg32 r;
    pcg_seed(&r, initstate, initseq);
    return pcg_next(&r);
}

/* the six uniforms one shade invocation uses: [0,1] bounce direction,
 * [2] Russian roulette, [3] emitter choice, [4,5] point on the emitter */
void orc_path_rng(uint32_t path, uint32_t bounce, uint32_t frame_seed, float out[6])
{
    pcg32 r;
    pcg_seed(&r, ((uint64_t)frame_seed << 32) | path, bounce);
    for (int k = 0; k < 6; ++k) out[k] = (float)(pcg_next(&r) >> 8) * (1.0f / 16777216.0f);
}

/* ---- sin / cos of a fraction of a turn -------------------------------------- */
/* theta = 2*pi*u = (pi/2)(q + f): quadrant q and fraction f from 4u (both
 * exact), then the Taylor series of sin and cos of x = f*pi/2 in [0, pi/2)
 * through x^13 and x^14 (truncation error < 1e-9), in Horner form on y = x^2. */
static void sincos_turns(float u, float* s, float* c)
{
    float q4 = u * 4.0f;
    float qf = floorf(q4);
    float f = q4 - qf;
    int q = ((int)qf) & 3;
    float x = f * 1.57079632679489662f;
    float y = x * x;
    float p = (float)(1.0 / 6227020800.0);      /* 1/13! */
    p = p * y - (float)(1.0 / 39916800.0);      /* 1/11! */
    p = p * y + (float)(1.0 / 362880.0);        /* 1/9!  */
    p = p * y - (float)(1.0 / 5040.0);          /* 1/7!  */
    p = p * y + (float)(1.0 / 120.0);           /* 1/5!  */
    p = p * y - (float)(1.0 / 6.0);             /* 1/3!  */
    p = p * y + 1.0f;
    float sx = x * p;
    float r = -(float)(1.0 / 87178291200.0);    /* 1/14! */
    r = r * y + (float)(1.0 / 479001600.0);     /* 1/12! */
    r = r * y - (float)(1.0 / 3628800.0);       /* 1/10! */
    r = r * y + (float)(1.0 / 40320.0);         /* 1/8!  */
    r = r * y - (float)(1.0 / 720.0);           /* 1/6!  */
    r = r * y + (float)(1.0 / 24.0);            /* 1/4!  */
    r = r * y - 0.5f;
    r = r * y + 1.0f;
    float cx = r;
    switch (q) {
    case 0: *s = sx; *c = cx; break;
    case 1: *s = cx; *c = -sx; break;
    case 2: *s = -sx; *c = -cx; break;
    default: *s = -cx; *c = sx; break;
    }
}

float orc_sin_turns(float u) { float s, c; sincos_turns(u, &s, &c); return s; }
float orc_cos_turns(float u) { float s, c; sincos_turns(u, &s, &c); return c; }

/* ---- small vector helpers (fixed evaluation order) ---------------------------- */
static float dot3(const float a[3], const float b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
static void normalize3(const float v[3], float o[3])
{
    float s = 1.0f / sqrtf(dot3(v, v));
    o[0] = v[0] * s; o[1] = v[1] * s; o[2] = v[2] * s;
}
static void cross3(const float a[3], const float b[3], float o[3])
{
    float x = a[1] * b[2] - a[2] * b[1];
    float y = a[2] * b[0] - a[0] * b[2];
    float z = a[0] * b[1] - a[1] * b[0];
    o[0] = x; o[1] = y; o[2] = z;
}
static void xform(const float m[16], const float v[3], float w, float o[3])
{
    float in[4] = { v[0], v[1], v[2], w }, r[4];
    orc_mat4_mul_vec4(m, in, r);
    o[0] = r[0]; o[1] = r[1]; o[2] = r[2];
}
/* interpolate (common.hlsl:212-216): v0 + b.x*(v1-v0) + b.y*(v2-v0) */
static float interp(float a, float b, float c, float bx, float by) { return (a + bx * (b - a)) + by * (c - a); }

/* cosine_weighted_hemisphere_sample with create_orthonormal_basis
 * (sample.hlsl:13-27, 45-54) */
void orc_cosine_sample(const float n[3], float r0, float r1, float out[3])
{
    float sin_theta = sqrtf(r0);
    float sp, cp;
    sincos_turns(r1, &sp, &cp);
    float v[3] = { sin_theta * cp, sqrtf(1.0f - r0), sin_theta * sp };
    float vn[3];
    normalize3(v, vn);
    float t0[3], t[3], b[3];
    if (fabsf(n[0]) > fabsf(n[2])) { t0[0] = -n[1]; t0[1] = n[0]; t0[2] = 0.0f; }
    else { t0[0] = 0.0f; t0[1] = -n[2]; t0[2] = n[1]; }
    normalize3(t0, t);
    cross3(n, t, b);
    for (int k = 0; k < 3; ++k) out[k] = (vn[0] * t[k] + vn[1] * n[k]) + vn[2] * b[k];
}

/* uniform_hemisphere_sample with create_orthonormal_basis (sample.hlsl:13-43) */
void orc_uniform_sample(const float n[3], float r0, float r1, float out[3])
{
    float sin_theta = sqrtf(1.0f - r0 * r0);
    float sp, cp;
    sincos_turns(r1, &sp, &cp);
    float v[3] = { sin_theta * cp, r0, sin_theta * sp };
    float vn[3];
    normalize3(v, vn);
    float t0[3], t[3], b[3];
    if (fabsf(n[0]) > fabsf(n[2])) { t0[0] = -n[1]; t0[1] = n[0]; t0[2] = 0.0f; }
    else { t0[0] = 0.0f; t0[1] = -n[2]; t0[2] = n[1]; }
    normalize3(t0, t);
    cross3(n, t, b);
    for (int k = 0; k < 3; ++k) out[k] = (vn[0] * t[k] + vn[1] * n[k]) + vn[2] * b[k];
}

/* bilinear, wrap addressing (the sampler of renderer.cpp:341-355, exact
 * weights); optionally reports the corners and fractions */
static void tex_sample(const orc_shading* sh, int32_t tex, float u, float v, float out[3], orc_vertex* rec)
{
    const float* t = sh->tex_rgba + sh->tex_off[tex];
    uint32_t w = sh->tex_w[tex], h = sh->tex_h[tex];
    float tx = u * (float)w - 0.5f, ty = v * (float)h - 0.5f;
    if (!(tx > -1e9f)) tx = -1e9f;
    if (!(tx < 1e9f)) tx = 1e9f;
    if (!(ty > -1e9f)) ty = -1e9f;
    if (!(ty < 1e9f)) ty = 1e9f;
    float fx0 = floorf(tx), fy0 = floorf(ty);
    float fx = tx - fx0, fy = ty - fy0;
    int64_t ix = (int64_t)fx0, iy = (int64_t)fy0;
    uint32_t x0 = (uint32_t)(((ix % (int64_t)w) + w) % w), y0 = (uint32_t)(((iy % (int64_t)h) + h) % h);
    uint32_t x1 = (x0 + 1u == w) ? 0u : x0 + 1u, y1 = (y0 + 1u == h) ? 0u : y0 + 1u;
    if (rec) {
        rec->texel[0] = y0 * w + x0; rec->texel[1] = y0 * w + x1;
        rec->texel[2] = y1 * w + x0; rec->texel[3] = y1 * w + x1;
        rec->fx = fx; rec->fy = fy;
    }
    for (int k = 0; k < 3; ++k) {
        float a = t[((size_t)y0 * w + x0) * 4 + k], b = t[((size_t)y0 * w + x1) * 4 + k];
        float c = t[((size_t)y1 * w + x0) * 4 + k], d = t[((size_t)y1 * w + x1) * 4 + k];
        float top = a + fx * (b - a), bot = c + fx * (d - c);
        out[k] = top + fy * (bot - top);
    }
}

/* ---- emitter table --------------------------------------------------------------- */
static void world_tri(const orc_shading* sh, uint32_t inst, uint32_t prim, float w[3][3])
{
    const orc_material_instance* m = &sh->mi[inst];
    const float* t = sh->shade_tris + 24 * ((size_t)sh->shade_off[m->mesh] + prim);
    for (int v = 0; v < 3; ++v) xform(m->local_to_world, &t[3 * v], 1.0f, w[v]);
}

uint32_t orc_emitters(const orc_shading* sh, uint32_t* ids, float* cdf, float* total_area)
{
    uint32_t n = 0;
    double total = 0.0;
    /* pass 1: count + total (double area of the float world triangle) */
    for (int pass = 0; pass < 2; ++pass) {
        double acc = 0.0;
        uint32_t k = 0;
        for (uint32_t i = 0; i < sh->n_inst; ++i) {
            const orc_material_instance* m = &sh->mi[i];
            if (!(m->emissive[0] > 0.0f || m->emissive[1] > 0.0f || m->emissive[2] > 0.0f)) continue;
            for (uint32_t t = 0; t < sh->tri_count[m->mesh]; ++t) {
                float w[3][3];
                world_tri(sh, i, t, w);
                double e1[3], e2[3];
                for (int c = 0; c < 3; ++c) { e1[c] = (double)w[1][c] - w[0][c]; e2[c] = (double)w[2][c] - w[0][c]; }
                double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
                double a = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
                if (!(a > 0.0)) continue;
                if (pass == 0) { total += a; ++n; continue; }
                acc += a;
                if (ids) { ids[2 * k] = i; ids[2 * k + 1] = t; }
                if (cdf) cdf[k] = (k + 1 == n) ? 1.0f : (float)(acc / total);
                ++k;
            }
        }
        if (pass == 0 && !ids && !cdf) break;
    }
    if (total_area) *total_area = (float)total;
    return n;
}

/* ---- one path --------------------------------------------------------------------- */
typedef struct {
    const orc_scene* s; const orc_shading* sh; const orc_view* v; const orc_render_params* p;
    uint32_t w, h; float* path_energy;
    const uint32_t* em_ids; const float* em_cdf; uint32_t n_em; float em_area;
    uint32_t first, count, stride; int rc;
    orc_vertex* rec;   /* optional: (max_bounces + 1) records per path */
} render_job;

static void primary_ray(const orc_view* v, uint32_t px, uint32_t py, orc_ray* r)
{
    float uvx = ((float)px + 0.5f) / v->render_dim[0];
    float uvy = ((float)py + 0.5f) / v->render_dim[1];
    uvy = 1.0f - uvy;
    float c4[4] = { 2.0f * uvx - 1.0f, 2.0f * uvy - 1.0f, 1.0f, 1.0f }, cv[4];
    orc_mat4_mul_vec4(v->clip_to_view, c4, cv);
    float dv[3];
    normalize3(cv, dv);
    float dw[3], ow[3], zero[3] = { 0.0f, 0.0f, 0.0f };
    xform(v->view_to_world, dv, 0.0f, dw);
    xform(v->view_to_world, zero, 1.0f, ow);
    for (int k = 0; k < 3; ++k) { r->origin[k] = ow[k] + dw[k] * RAY_MIN_T; r->dir[k] = dw[k]; }
    r->tmin = RAY_MIN_T;
    r->tmax = FLT_MAX;
}

/* One path; rec (max_bounces + 1 entries, may be NULL) receives the vertex
 * records for the FP64 evaluator. */
uint32_t orc_env_texel(const orc_shading* sh, const float d[3])
{
    const uint32_t w = sh->tex_w[sh->env_tex], h = sh->tex_h[sh->env_tex];
    const float u = atan2f(d[2], d[0]) * 0.1591f + 0.5f;
    const float v = asinf(-d[1]) * 0.3183f + 0.5f;
    const float fx = u * (float)w, fy = v * (float)h;
    uint32_t x = fx >= 1.0f ? (uint32_t)fx : 0u, y = fy >= 1.0f ? (uint32_t)fy : 0u;
    if (x >= w) x = w - 1;
    if (y >= h) y = h - 1;
    return y * w + x;
}

/* murmur_mix (common.hlsl:276-284): MurmurHash3's 32-bit finaliser */
uint32_t orc_murmur_mix(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

/* int_to_color (common.hlsl:286-295) */
static void int_to_color(uint32_t v, float out[3])
{
    const uint32_t h = orc_murmur_mix(v);
    const float s = 1.0f / 255.0f;
    out[0] = (float)(h & 255u) * s;
    out[1] = (float)((h >> 8) & 255u) * s;
    out[2] = (float)((h >> 16) & 255u) * s;
}

/* The debug view modes (shade.hlsl:140-153) for a camera ray's hit.  A miss
 * sees the reference's zeroed records: its ids, bary (0, 0) and t, zero
 * geometry and material values, normalize(0) as its world normal; the
 * material normal and metallic-roughness views of a miss are 0. */
static void view_color(const render_job* J, const orc_hit* hit, float out[3])
{
    const orc_shading* sh = J->sh;
    const uint32_t vm = J->p->view_mode;
    out[0] = out[1] = out[2] = 0.0f;
    if (vm == 1) { int_to_color(hit->instance_idx, out); return; }
    if (vm == 2) { int_to_color(hit->primitive_idx, out); return; }
    if (vm == 3) { out[0] = hit->bary[0]; out[1] = hit->bary[1]; return; }
    if (vm == 11) { out[0] = out[1] = out[2] = hit->t / J->v->far_plane; return; }
    if (hit->instance_idx == ~0u) {
        if (vm == 10) {
            const float z[3] = { 0.0f, 0.0f, 0.0f };
            float n[3];
            normalize3(z, n);
            for (int k = 0; k < 3; ++k) out[k] = fabsf(n[k]);
        }
        return;
    }
    const orc_material_instance* m = &sh->mi[hit->instance_idx];
    const float* t = sh->shade_tris + 24 * ((size_t)sh->shade_off[m->mesh] + hit->primitive_idx);
    const float bx = hit->bary[0], by = hit->bary[1];
    const float u = interp(t[18], t[20], t[22], bx, by), v = interp(t[19], t[21], t[23], bx, by);
    float ln[3], n0[3], n[3], c[3];
    for (int k = 0; k < 3; ++k) ln[k] = interp(t[9 + k], t[12 + k], t[15 + k], bx, by);
    switch (vm) {
    case 4: for (int k = 0; k < 3; ++k) out[k] = fabsf(ln[k]); break;
    case 5: out[0] = u; out[1] = v; break;
    case 6:
        for (int k = 0; k < 3; ++k) out[k] = m->base_color[k];
        if (m->base_color_tex >= 0) {
            tex_sample(sh, m->base_color_tex, u, v, c, NULL);
            for (int k = 0; k < 3; ++k) out[k] = out[k] * c[k];
        }
        break;
    case 7:   /* default normal texture: RGBA8 (127, 127, 255, 255), renderer.cpp:256-262 */
        c[0] = 127.0f / 255.0f; c[1] = 127.0f / 255.0f; c[2] = 1.0f;
        if (m->normal_tex >= 0) tex_sample(sh, m->normal_tex, u, v, c, NULL);
        for (int k = 0; k < 3; ++k) out[k] = fabsf(c[k]);
        break;
    case 8:   /* (0, g, b); default texture RG8 (255, 255), renderer.cpp:264-269 */
        c[0] = 1.0f; c[1] = 1.0f; c[2] = 0.0f;
        if (m->metal_rough_tex >= 0) tex_sample(sh, m->metal_rough_tex, u, v, c, NULL);
        out[1] = c[1]; out[2] = c[2];
        break;
    case 9:
        c[0] = c[1] = c[2] = 1.0f;
        if (m->emissive_tex >= 0) tex_sample(sh, m->emissive_tex, u, v, c, NULL);
        for (int k = 0; k < 3; ++k) out[k] = m->emissive[k] * c[k];
        break;
    case 10:
        xform(m->local_to_world, ln, 0.0f, n0);
        normalize3(n0, n);
        for (int k = 0; k < 3; ++k) out[k] = fabsf(n[k]);
        break;
    default: break;
    }
}

static int trace_path(render_job* J, uint32_t path, float energy[3], orc_vertex* rec)
{
    const orc_render_params* P = J->p;
    const orc_shading* sh = J->sh;
    uint32_t npix = J->w * J->h;
    uint32_t pixel = path % npix;
    path += P->path_base;
    orc_ray ray;
    primary_ray(J->v, pixel % J->w, pixel / J->w, &ray);
    float thr[3] = { 1.0f, 1.0f, 1.0f };
    energy[0] = energy[1] = energy[2] = 0.0f;
    if (rec) memset(rec, 0, sizeof(orc_vertex) * (P->max_bounces + 1));
    for (uint32_t d = 0; d <= P->max_bounces; ++d) {
        orc_vertex* vr = rec ? &rec[d] : NULL;
        orc_hit hit;
        int rc = orc_trace_one(J->s, &ray, &hit, 0);
        if (rc) return rc;
        if (P->view_mode != 0) {   /* a debug view replaces the radiance and ends the path */
            view_color(J, &hit, energy);
            break;
        }
        float u[6];
        orc_path_rng(path, d, P->frame_seed, u);
        if (hit.instance_idx == ~0u) {
            if (sh->env_tex >= 0) {
                /* throughput * hdr_env_strength * env (shade.hlsl:62-63) */
                const uint32_t ti = orc_env_texel(sh, ray.dir);
                const float* c = sh->tex_rgba + sh->tex_off[sh->env_tex] + 4 * (size_t)ti;
                for (int k = 0; k < 3; ++k) energy[k] = energy[k] + (thr[k] * sh->env_strength) * c[k];
                if (vr) vr->texel[0] = ti;
            } else {
                for (int k = 0; k < 3; ++k) energy[k] = energy[k] + thr[k] * sh->env[k];
            }
            if (vr) vr->kind = ORC_V_MISS;
            break;
        }
        if (vr) { vr->inst = hit.instance_idx; vr->tex = -1; }
        const orc_material_instance* m = &sh->mi[hit.instance_idx];
        const float* t = sh->shade_tris + 24 * ((size_t)sh->shade_off[m->mesh] + hit.primitive_idx);
        float bx = hit.bary[0], by = hit.bary[1];
        float lp[3], ln[3], uv[2], P3[3], n0[3], N[3];
        for (int k = 0; k < 3; ++k) {
            lp[k] = interp(t[k], t[3 + k], t[6 + k], bx, by);
            ln[k] = interp(t[9 + k], t[12 + k], t[15 + k], bx, by);
        }
        uv[0] = interp(t[18], t[20], t[22], bx, by);
        uv[1] = interp(t[19], t[21], t[23], bx, by);
        xform(m->local_to_world, lp, 1.0f, P3);
        xform(m->local_to_world, ln, 0.0f, n0);
        normalize3(n0, N);
        float albedo[3] = { m->base_color[0], m->base_color[1], m->base_color[2] };
        if (m->base_color_tex >= 0) {
            float tc[3];
            tex_sample(sh, m->base_color_tex, uv[0], uv[1], tc, vr);
            for (int k = 0; k < 3; ++k) albedo[k] = albedo[k] * tc[k];
            if (vr) vr->tex = m->base_color_tex;
        }
        /* emission = strength * factor * emissive texel (sample_material, material.hlsl:35-37) */
        float etc[3] = { 1.0f, 1.0f, 1.0f }, em[3];
        if (m->emissive_tex >= 0) tex_sample(sh, m->emissive_tex, uv[0], uv[1], etc, NULL);
        for (int k = 0; k < 3; ++k) em[k] = m->emissive[k] * etc[k];
        if (em[0] != 0.0f || em[1] != 0.0f || em[2] != 0.0f) {
            if (!P->nee || d == 0)
                for (int k = 0; k < 3; ++k) energy[k] = energy[k] + thr[k] * em[k];
            if (vr) {
                vr->kind = ORC_V_EMIT;
                vr->emit_counted = (!P->nee || d == 0) ? 1u : 0u;
                for (int k = 0; k < 3; ++k) vr->em_tex[k] = etc[k];
            }
            break;
        }
        if (vr) vr->kind = ORC_V_DIFFUSE;
        if (P->nee && J->n_em > 0 && d < P->max_bounces) {
            /* emitter by area: first k with cdf[k] > u[3] */
            uint32_t lo = 0, hi = J->n_em - 1;
            while (lo < hi) {
                uint32_t mid = (lo + hi) >> 1;
                if (J->em_cdf[mid] > u[3]) hi = mid; else lo = mid + 1;
            }
            uint32_t li = J->em_ids[2 * lo], lt = J->em_ids[2 * lo + 1];
            float W[3][3];
            world_tri(sh, li, lt, W);
            float su = sqrtf(u[4]);
            float bv = su * (1.0f - u[5]), bw = su * u[5];
            float PL[3], dl[3], e1[3], e2[3], nl0[3], nl[3], wi[3];
            for (int k = 0; k < 3; ++k) {
                PL[k] = interp(W[0][k], W[1][k], W[2][k], bv, bw);
                dl[k] = PL[k] - P3[k];
                e1[k] = W[1][k] - W[0][k];
                e2[k] = W[2][k] - W[0][k];
            }
            float dist2 = dot3(dl, dl);
            float dist = sqrtf(dist2);
            float inv_d = 1.0f / dist;
            for (int k = 0; k < 3; ++k) wi[k] = dl[k] * inv_d;
            cross3(e1, e2, nl0);
            normalize3(nl0, nl);
            float cos_s = dot3(N, wi);
            float cos_l = -dot3(nl, wi);
            if (cos_s > 0.0f && cos_l > 0.0f && dist2 > 0.0f) {
                float G = (cos_s * cos_l) / dist2;
                float wgt = G * J->em_area;
                /* the light's emission at the sampled point, with its emissive texture */
                float ltc[3] = { 1.0f, 1.0f, 1.0f }, Le[3];
                const orc_material_instance* lm = &sh->mi[li];
                if (lm->emissive_tex >= 0) {
                    const float* lt6 = sh->shade_tris + 24 * ((size_t)sh->shade_off[lm->mesh] + lt);
                    const float lu = interp(lt6[18], lt6[20], lt6[22], bv, bw);
                    const float lv = interp(lt6[19], lt6[21], lt6[23], bv, bw);
                    tex_sample(sh, lm->emissive_tex, lu, lv, ltc, NULL);
                }
                for (int k = 0; k < 3; ++k) Le[k] = lm->emissive[k] * ltc[k];
                float c[3];
                for (int k = 0; k < 3; ++k) c[k] = (((thr[k] * albedo[k]) * INV_PI) * Le[k]) * wgt;
                orc_ray sr;
                for (int k = 0; k < 3; ++k) { sr.origin[k] = P3[k] + wi[k] * RAY_MIN_T; sr.dir[k] = wi[k]; }
                sr.tmin = RAY_MIN_T;
                sr.tmax = dist * 0.9999f;
                orc_hit sh_hit;
                rc = orc_trace_one(J->s, &sr, &sh_hit, 1);
                if (rc) return rc;
                if (sh_hit.instance_idx == ~0u) {
                    for (int k = 0; k < 3; ++k) energy[k] = energy[k] + c[k];
                    if (vr) {
                        vr->nee = 1; vr->light_inst = li; vr->nee_wgt = wgt;
                        for (int k = 0; k < 3; ++k) vr->nee_tex[k] = ltc[k];
                    }
                }
            }
        }
        /* shade.hlsl:108-135: cosine (default) or uniform hemisphere sample */
        float L[3];
        if (P->uniform_diffuse) orc_uniform_sample(N, u[0], u[1], L);
        else orc_cosine_sample(N, u[0], u[1], L);
        float NoL = dot3(N, L);
        if (NoL < 0.0f) NoL = 0.0f;
        if (!(NoL > 0.0f)) break;
        float pdf = P->uniform_diffuse ? INV_TWO_PI : NoL * INV_PI;
        float inv_pdf = 1.0f / pdf;
        for (int k = 0; k < 3; ++k) thr[k] = thr[k] * ((NoL * (albedo[k] * INV_PI)) * inv_pdf);
        if (vr) { vr->noL = NoL; vr->inv_pdf = inv_pdf; }
        if (d >= P->rr_start) {
            if (u[2] >= P->rr_survival) break;
            for (int k = 0; k < 3; ++k) thr[k] = thr[k] / P->rr_survival;
            if (vr) vr->rr = 1;
        }
        if (d >= P->max_bounces) break;
        for (int k = 0; k < 3; ++k) { ray.origin[k] = P3[k] + L[k] * RAY_MIN_T; ray.dir[k] = L[k]; }
        ray.tmin = RAY_MIN_T;
        ray.tmax = FLT_MAX;
    }
    return ORC_OK;
}

static void* render_worker(void* arg)
{
    render_job* J = (render_job*)arg;
    J->rc = ORC_OK;
    for (uint32_t p = J->first; p < J->count; p += J->stride) {
        orc_vertex* rec = J->rec ? J->rec + (size_t)p * (J->p->max_bounces + 1) : NULL;
        int rc = trace_path(J, p, &J->path_energy[3 * (size_t)p], rec);
        if (rc) J->rc = rc;
    }
    return NULL;
}

static int run_paths(const orc_scene* s, const orc_shading* sh, const orc_view* v, uint32_t width, uint32_t height,
                     const orc_render_params* p, float* pe, orc_vertex* rec, int nthreads)
{
    uint32_t npath = width * height * p->spp;
    uint32_t n_em = orc_emitters(sh, NULL, NULL, NULL);
    uint32_t* ids = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (n_em ? n_em : 1));
    float* cdf = (float*)malloc(sizeof(float) * (n_em ? n_em : 1));
    if (!ids || !cdf) { free(ids); free(cdf); return ORC_ECAP; }
    float area = 0.0f;
    orc_emitters(sh, ids, cdf, &area);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    render_job jobs[64];
    pthread_t th[64];
    for (int t = 0; t < nthreads; ++t) {
        render_job j = { s, sh, v, p, width, height, pe, ids, cdf, n_em, area, (uint32_t)t, npath, (uint32_t)nthreads, 0, rec };
        jobs[t] = j;
    }
    int rc = ORC_OK;
    if (nthreads == 1) {
        render_worker(&jobs[0]);
        rc = jobs[0].rc;
    } else {
        for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, render_worker, &jobs[t]);
        for (int t = 0; t < nthreads; ++t) { pthread_join(th[t], NULL); if (jobs[t].rc) rc = jobs[t].rc; }
    }
    free(ids); free(cdf);
    return rc;
}

int orc_record_paths(const orc_scene* s, const orc_shading* sh, const orc_view* v, uint32_t width, uint32_t height,
                     const orc_render_params* p, orc_vertex* rec, int nthreads)
{
    if (!s || !sh || !v || !p || !rec || width == 0 || height == 0 || p->spp == 0) return ORC_EINVAL;
    float* pe = (float*)malloc(sizeof(float) * 3 * (size_t)width * height * p->spp);
    if (!pe) return ORC_ECAP;
    int rc = run_paths(s, sh, v, width, height, p, pe, rec, nthreads);
    free(pe);
    return rc;
}

int orc_render(const orc_scene* s, const orc_shading* sh, const orc_view* v, uint32_t width, uint32_t height,
               const orc_render_params* p, float* energy_rgba, int nthreads)
{
    if (!s || !sh || !v || !p || !energy_rgba || width == 0 || height == 0 || p->spp == 0) return ORC_EINVAL;
    uint32_t npix = width * height;
    uint32_t npath = npix * p->spp;
    float* pe = (float*)malloc(sizeof(float) * 3 * (size_t)npath);
    if (!pe) return ORC_ECAP;
    int rc = run_paths(s, sh, v, width, height, p, pe, NULL, nthreads);
    /* resolve: mean over the spp paths of each pixel, samples in order */
    for (uint32_t px = 0; px < npix; ++px) {
        float acc[3] = { 0.0f, 0.0f, 0.0f };
        for (uint32_t sidx = 0; sidx < p->spp; ++sidx)
            for (int k = 0; k < 3; ++k) acc[k] = acc[k] + pe[3 * ((size_t)sidx * npix + px) + k];
        for (int k = 0; k < 3; ++k) energy_rgba[4 * (size_t)px + k] = acc[k] / (float)p->spp;
        energy_rgba[4 * (size_t)px + 3] = 1.0f;
    }
    free(pe);
    return rc;
}

