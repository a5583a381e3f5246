/*
 * Oracle: ray/triangle and ray/AABB tests, TLAS->BLAS stack traversal,
 * occlusion queries and brute force.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

#include <float.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define RAY_MAX_T FLT_MAX
#define INTERSECT_EPSILON 1e-8f   /* common.hlsl:8,12: the float literal 1e-8 */
#define STACK_SIZE 64             /* accelstruct.hlsl:64,160 */

/* HLSL min/max: IEEE minNum/maxNum (a NaN operand yields the other one). */
static inline float hmin(float a, float b) { return (a < b || b != b) ? a : b; }
static inline float hmax(float a, float b) { return (a > b || b != b) ? a : b; }

static inline float dot3(const float a[3], const float b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
static inline void cross3(const float a[3], const float b[3], float o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

/* intersect_ray_triangle, intersect.hlsl:7-51 (back-face culling on). */
int orc_intersect_triangle(const float v0[3], const float v1[3], const float v2[3],
                           const float o[3], const float d[3], float* t_inout, float bary[2])
{
    float e1[3] = { v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2] };
    float e2[3] = { v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2] };
    float p[3];
    cross3(d, e2, p);
    float det = dot3(e1, p);
    if (det < INTERSECT_EPSILON) return 0;
    float inv = 1.0f / det;
    float s[3] = { o[0] - v0[0], o[1] - v0[1], o[2] - v0[2] };
    float v = dot3(s, p) * inv;
    if (v < 0.0f || v > 1.0f) return 0;
    float q[3];
    cross3(s, e1, q);
    float w = dot3(d, q) * inv;
    if (w < 0.0f || v + w > 1.0f) return 0;
    float t = dot3(e2, q) * inv;
    if (t < 0.0f || t >= *t_inout) return 0;
    *t_inout = t;
    bary[0] = v;
    bary[1] = w;
    return 1;
}

/* intersect_ray_aabb, intersect.hlsl:53-76. */
float orc_intersect_aabb(const float bmin[3], const float bmax[3], const float o[3], const float inv[3], float t)
{
    float tx1 = (bmin[0] - o[0]) * inv[0];
    float tx2 = (bmax[0] - o[0]) * inv[0];
    float tmin = hmin(tx1, tx2);
    float tmax = hmax(tx1, tx2);
    float ty1 = (bmin[1] - o[1]) * inv[1];
    float ty2 = (bmax[1] - o[1]) * inv[1];
    tmin = hmax(tmin, hmin(ty1, ty2));
    tmax = hmin(tmax, hmax(ty1, ty2));
    float tz1 = (bmin[2] - o[2]) * inv[2];
    float tz2 = (bmax[2] - o[2]) * inv[2];
    tmin = hmax(tmin, hmin(tz1, tz2));
    tmax = hmin(tmax, hmax(tz1, tz2));
    if (tmax >= tmin && tmin < t && tmax > 0.0f) return tmin;
    return RAY_MAX_T;
}

/* world -> object ray, accelstruct.hlsl:139-142, with mul(v, M) taken as
 * the glm product M*v in glm's association order. */
static void to_local(const float m[16], const float o[3], const float d[3], float lo[3], float ld[3], float linv[3])
{
    float ow[4] = { o[0], o[1], o[2], 1.0f }, dw[4] = { d[0], d[1], d[2], 0.0f };
    float a[4], b[4];
    orc_mat4_mul_vec4(m, ow, a);
    orc_mat4_mul_vec4(m, dw, b);
    for (int k = 0; k < 3; ++k) { lo[k] = a[k]; ld[k] = b[k]; linv[k] = 1.0f / ld[k]; }
}

typedef struct { float t; uint32_t prim; float bary[2]; } blas_hit;

/* trace_ray_bvh_local, accelstruct.hlsl:58-131.  any != 0 stops at the first
 * accepted triangle (occlusion). */
static int trace_blas(const orc_scene* s, uint32_t blas, const float o[3], const float d[3], const float inv[3],
                      float* t, uint32_t* prim, float bary[2], orc_counters* c, int any, int* overflow)
{
    const orc_node* nodes = s->nodes + s->node_off[blas];
    const orc_tri* tris = s->tris + s->tri_off[blas];
    const uint32_t* idx = s->idx + s->tri_off[blas];
    int has = 0;
    uint32_t stack[STACK_SIZE];
    uint32_t sp = 0;
    uint32_t ni = 0;
    if (c) c->blas_nodes++;
    for (;;) {
        const orc_node* n = &nodes[ni];
        if (n->prim_count > 0) {
            for (uint32_t i = n->left_first; i < n->left_first + n->prim_count; ++i) {
                uint32_t ti = idx[i];
                const orc_tri* tr = &tris[ti];
                if (c) c->tri_tests++;
                if (orc_intersect_triangle(tr->p0, tr->p1, tr->p2, o, d, t, bary)) {
                    *prim = ti;
                    has = 1;
                    if (any) return 1;
                }
            }
            if (sp == 0) break;
            ni = stack[--sp];
            continue;
        }
        uint32_t li = n->left_first, ri = n->left_first + 1;
        if (c) c->blas_nodes += 2;
        float dl = orc_intersect_aabb(nodes[li].aabb_min, nodes[li].aabb_max, o, inv, *t);
        float dr = orc_intersect_aabb(nodes[ri].aabb_min, nodes[ri].aabb_max, o, inv, *t);
        if (dl > dr) { float tf = dl; dl = dr; dr = tf; uint32_t tu = li; li = ri; ri = tu; }
        if (dl == RAY_MAX_T) {
            if (sp == 0) break;
            ni = stack[--sp];
        } else {
            ni = li;
            if (dr != RAY_MAX_T) {
                if (sp == STACK_SIZE) { *overflow = 1; return has; }
                stack[sp++] = ri;
            }
        }
    }
    return has;
}

/* trace_ray_tlas, accelstruct.hlsl:151-218 */
static int trace_one(const orc_scene* s, const orc_ray* r, orc_hit* h, orc_counters* c, int any)
{
    h->instance_idx = ~0u; h->primitive_idx = ~0u; h->t = RAY_MAX_T; h->bary[0] = 0.0f; h->bary[1] = 0.0f;
    if (c) memset(c, 0, sizeof(*c));
    float t = r->tmax;
    float inv[3] = { 1.0f / r->dir[0], 1.0f / r->dir[1], 1.0f / r->dir[2] };
    int overflow = 0;
    const orc_tnode* tn = s->tnodes;
    if (c) c->tlas_nodes++;
    if (orc_intersect_aabb(tn[0].aabb_min, tn[0].aabb_max, r->origin, inv, t) == RAY_MAX_T) return ORC_OK;
    uint32_t stack[STACK_SIZE];
    uint32_t sp = 0;
    uint32_t ni = 0;
    for (;;) {
        const orc_tnode* n = &tn[ni];
        if (n->left_right == 0) {
            const orc_instance* in = &s->inst[n->instance_idx];
            if (c) c->instances++;
            float lo[3], ld[3], linv[3];
            to_local(in->world_to_local, r->origin, r->dir, lo, ld, linv);
            uint32_t prim = 0;
            float bary[2] = { h->bary[0], h->bary[1] };
            if (trace_blas(s, in->bvh_index, lo, ld, linv, &t, &prim, bary, c, any, &overflow)) {
                h->instance_idx = n->instance_idx;
                h->primitive_idx = prim;
                h->t = t;
                h->bary[0] = bary[0];
                h->bary[1] = bary[1];
                if (any) return ORC_OK;
            }
            if (overflow) return ORC_ESTACK;
            if (sp == 0) break;
            ni = stack[--sp];
            continue;
        }
        uint32_t li = n->left_right >> 16, ri = n->left_right & 0xFFFFu;
        if (c) c->tlas_nodes += 2;
        float dl = orc_intersect_aabb(tn[li].aabb_min, tn[li].aabb_max, r->origin, inv, t);
        float dr = orc_intersect_aabb(tn[ri].aabb_min, tn[ri].aabb_max, r->origin, inv, t);
        if (dl > dr) { float tf = dl; dl = dr; dr = tf; uint32_t tu = li; li = ri; ri = tu; }
        if (dl == RAY_MAX_T) {
            if (sp == 0) break;
            ni = stack[--sp];
        } else {
            ni = li;
            if (dr != RAY_MAX_T) {
                if (sp == STACK_SIZE) return ORC_ESTACK;
                stack[sp++] = ri;
            }
        }
    }
    return ORC_OK;
}

int orc_trace_one(const orc_scene* s, const orc_ray* r, orc_hit* h, int any)
{
    return trace_one(s, r, h, NULL, any);
}

typedef struct {
    const orc_scene* s; const orc_ray* rays; orc_hit* hits; orc_counters* cnt; uint32_t* vis;
    uint32_t n; int tid, nt; int any; int rc;
} job;

static void* run_job(void* p)
{
    job* j = (job*)p;
    j->rc = ORC_OK;
    /* interleaved 64-ray tiles so the threads see similar ray mixes */
    for (uint32_t base = (uint32_t)j->tid * 64u; base < j->n; base += (uint32_t)j->nt * 64u) {
        uint32_t end = base + 64u < j->n ? base + 64u : j->n;
        for (uint32_t i = base; i < end; ++i) {
            orc_hit h;
            int rc = trace_one(j->s, &j->rays[i], &h, j->cnt ? &j->cnt[i] : NULL, j->any);
            if (rc != ORC_OK) j->rc = rc;
            if (j->any) j->vis[i] = (h.instance_idx == ~0u) ? 1u : 0u;
            else j->hits[i] = h;
        }
    }
    return NULL;
}

static int run_all(const orc_scene* s, const orc_ray* rays, uint32_t n, orc_hit* hits, orc_counters* cnt,
                   uint32_t* vis, int any, int nthreads)
{
    if (!s || !rays || (!hits && !vis)) return ORC_EINVAL;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    job jobs[256];
    pthread_t th[256];
    for (int t = 0; t < nthreads; ++t) {
        jobs[t] = (job){ s, rays, hits, cnt, vis, n, t, nthreads, any, ORC_OK };
    }
    if (nthreads == 1) { run_job(&jobs[0]); return jobs[0].rc; }
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, run_job, &jobs[t]);
    int rc = ORC_OK;
    for (int t = 0; t < nthreads; ++t) { pthread_join(th[t], NULL); if (jobs[t].rc) rc = jobs[t].rc; }
    return rc;
}

int orc_trace_closest(const orc_scene* s, const orc_ray* rays, uint32_t n, orc_hit* hits,
                      orc_counters* counters, int nthreads)
{
    return run_all(s, rays, n, hits, counters, NULL, 0, nthreads);
}

int orc_trace_any(const orc_scene* s, const orc_ray* rays, uint32_t n, uint32_t* visible, int nthreads)
{
    return run_all(s, rays, n, NULL, NULL, visible, 1, nthreads);
}

int orc_brute_closest(const orc_scene* s, const orc_ray* rays, uint32_t n, orc_hit* hits)
{
    if (!s || !rays || !hits) return ORC_EINVAL;
    for (uint32_t i = 0; i < n; ++i) {
        const orc_ray* r = &rays[i];
        orc_hit* h = &hits[i];
        h->instance_idx = ~0u; h->primitive_idx = ~0u; h->t = RAY_MAX_T; h->bary[0] = h->bary[1] = 0.0f;
        float t = r->tmax;
        for (uint32_t k = 0; k < s->n_inst; ++k) {
            const orc_instance* in = &s->inst[k];
            float lo[3], ld[3], linv[3];
            to_local(in->world_to_local, r->origin, r->dir, lo, ld, linv);
            uint32_t b = in->bvh_index;
            uint32_t ntri = s->tri_off[b + 1] - s->tri_off[b];
            const orc_tri* tris = s->tris + s->tri_off[b];
            for (uint32_t ti = 0; ti < ntri; ++ti) {
                float bary[2];
                if (orc_intersect_triangle(tris[ti].p0, tris[ti].p1, tris[ti].p2, lo, ld, &t, bary)) {
                    h->instance_idx = k; h->primitive_idx = ti; h->t = t; h->bary[0] = bary[0]; h->bary[1] = bary[1];
                }
            }
        }
    }
    return ORC_OK;
}
