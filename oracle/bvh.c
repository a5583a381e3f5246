/*
 * Oracle: BLAS / TLAS builders and the glm matrix helpers they depend on.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

#include <float.h>
#include <stdlib.h>
#include <string.h>

/* glm::min / glm::max on floats and ints (extern/glm/detail/func_common.inl:17-35). */
static inline float gmin(float x, float y) { return (y < x) ? y : x; }
static inline float gmax(float x, float y) { return (x < y) ? y : x; }
/* Bin of a centroid: min(B-1, (int32_t)((c - lo) * scale)) as in
 * bvh_builder.cpp:194-195,130-131.  When the centroid extent is so small
 * that scale overflows, the product is inf or NaN and the reference's
 * float->int conversion is undefined; here it saturates (inf -> B-1,
 * NaN -> 0).  For finite products the value is the reference's. */
static inline int32_t bin_of(float c, float lo, float scale, uint32_t bins)
{
    float f = (c - lo) * scale;
    if (!(f >= 0.0f)) return 0;
    if (f >= (float)bins) return (int32_t)bins - 1;
    int32_t b = (int32_t)f;
    return b < (int32_t)bins - 1 ? b : (int32_t)bins - 1;
}

/* as_util::get_aabb_volume (as_util.h:7-11): half surface area. */
static inline float half_area(const float mn[3], const float mx[3])
{
    float ex = mx[0] - mn[0], ey = mx[1] - mn[1], ez = mx[2] - mn[2];
    return ex * ey + ey * ez + ez * ex;
}

/* ------------------------------------------------------------------------ */
/* BLAS builder: bvh_builder.cpp:6-50, 71-98, 100-168, 170-255             */
/* ------------------------------------------------------------------------ */

typedef struct {
    orc_node* nodes;
    uint32_t node_at;
    const orc_tri* tris;
    uint32_t* idx;
    float* cen;            /* 3 floats per triangle */
    uint32_t bins;
    int single_prim;
    /* scratch bins */
    float* bin_min; float* bin_max; uint32_t* bin_cnt;
    float* left_area; float* right_area;
} blas_state;

/* calc_node_min_max, bvh_builder.cpp:71-93 */
static void calc_node_min_max(blas_state* b, orc_node* node, float cmin[3], float cmax[3])
{
    for (int k = 0; k < 3; ++k) {
        node->aabb_min[k] = FLT_MAX; node->aabb_max[k] = -FLT_MAX;
        cmin[k] = FLT_MAX; cmax[k] = -FLT_MAX;
    }
    for (uint32_t i = node->left_first; i < node->left_first + node->prim_count; ++i) {
        const orc_tri* t = &b->tris[b->idx[i]];
        float tmn[3], tmx[3];
        for (int k = 0; k < 3; ++k) {
            /* get_triangle_min_max, bvh_builder.cpp:257-264 */
            tmn[k] = gmin(t->p0[k], t->p1[k]);
            tmx[k] = gmax(t->p0[k], t->p1[k]);
            tmn[k] = gmin(tmn[k], t->p2[k]);
            tmx[k] = gmax(tmx[k], t->p2[k]);
            node->aabb_min[k] = gmin(node->aabb_min[k], tmn[k]);
            node->aabb_max[k] = gmax(node->aabb_max[k], tmx[k]);
        }
        const float* c = &b->cen[3 * (size_t)b->idx[i]];
        for (int k = 0; k < 3; ++k) {
            cmin[k] = gmin(cmin[k], c[k]);
            cmax[k] = gmax(cmax[k], c[k]);
        }
    }
}

/* find_best_split_plane, bvh_builder.cpp:170-250 */
static float find_best_split_plane(blas_state* b, const orc_node* node, const float cmin[3], const float cmax[3],
                                   uint32_t* out_axis, uint32_t* out_split)
{
    float best = FLT_MAX;
    const uint32_t B = b->bins;
    for (uint32_t axis = 0; axis < 3; ++axis) {
        float lo = cmin[axis], hi = cmax[axis];
        if (lo == hi) continue;
        for (uint32_t i = 0; i < B; ++i) {
            b->bin_cnt[i] = 0;
            for (int k = 0; k < 3; ++k) { b->bin_min[3 * i + k] = FLT_MAX; b->bin_max[3 * i + k] = -FLT_MAX; }
        }
        float scale = (float)B / (hi - lo);
        for (uint32_t i = node->left_first; i < node->left_first + node->prim_count; ++i) {
            const orc_tri* t = &b->tris[b->idx[i]];
            const float* c = &b->cen[3 * (size_t)b->idx[i]];
            int32_t bi = bin_of(c[axis], lo, scale, B);
            b->bin_cnt[bi]++;
            const float* ps[3] = { t->p0, t->p1, t->p2 };
            for (int v = 0; v < 3; ++v)
                for (int k = 0; k < 3; ++k) {
                    b->bin_min[3 * bi + k] = gmin(b->bin_min[3 * bi + k], ps[v][k]);
                    b->bin_max[3 * bi + k] = gmax(b->bin_max[3 * bi + k], ps[v][k]);
                }
        }
        float lmn[3] = { FLT_MAX, FLT_MAX, FLT_MAX }, lmx[3] = { -FLT_MAX, -FLT_MAX, -FLT_MAX };
        float rmn[3] = { FLT_MAX, FLT_MAX, FLT_MAX }, rmx[3] = { -FLT_MAX, -FLT_MAX, -FLT_MAX };
        uint32_t lsum = 0, rsum = 0;
        for (uint32_t i = 0; i < B - 1; ++i) {
            lsum += b->bin_cnt[i];
            for (int k = 0; k < 3; ++k) {
                lmn[k] = gmin(lmn[k], b->bin_min[3 * i + k]);
                lmx[k] = gmax(lmx[k], b->bin_max[3 * i + k]);
            }
            b->left_area[i] = (float)lsum * half_area(lmn, lmx);
            uint32_t r = B - 1 - i;
            rsum += b->bin_cnt[r];
            for (int k = 0; k < 3; ++k) {
                rmn[k] = gmin(rmn[k], b->bin_min[3 * r + k]);
                rmx[k] = gmax(rmx[k], b->bin_max[3 * r + k]);
            }
            b->right_area[B - 2 - i] = (float)rsum * half_area(rmn, rmx);
        }
        for (uint32_t i = 0; i < B - 1; ++i) {
            float cost = b->left_area[i] + b->right_area[i];
            if (cost < best) { *out_axis = axis; *out_split = i + 1; best = cost; }
        }
    }
    return best;
}

typedef struct { uint32_t node; } work_item;

/* subdivide_node, bvh_builder.cpp:100-168, run with an explicit LIFO stack
 * that visits nodes in the same depth-first pre-order as the recursion. */
static int subdivide_all(blas_state* b, uint32_t root)
{
    uint32_t cap = 64, sp = 0;
    uint32_t* stack = (uint32_t*)malloc(cap * sizeof(uint32_t));
    if (!stack) return ORC_ECAP;
    stack[sp++] = root;
    /* the root's min/max was computed by the caller; children get theirs
     * when popped, as in the recursive version */
    int first = 1;
    while (sp) {
        uint32_t ni = stack[--sp];
        orc_node* node = &b->nodes[ni];
        float cmin[3], cmax[3];
        calc_node_min_max(b, node, cmin, cmax);
        (void)first; first = 0;

        uint32_t axis = 0, split = 0;
        float split_cost = find_best_split_plane(b, node, cmin, cmax, &axis, &split);
        if (b->single_prim) {
            if (node->prim_count == 1) continue;
        } else {
            float parent_cost = (float)node->prim_count * half_area(node->aabb_min, node->aabb_max);
            if (split_cost >= parent_cost) continue;
        }
        if (split_cost == FLT_MAX) {
            /* single-prim mode with no splittable axis: the reference's
             * partition then sends every primitive left (x86 float->int of
             * NaN is INT_MIN) and the node stays a leaf */
            continue;
        }

        int32_t i = (int32_t)node->left_first;
        int32_t j = i + (int32_t)node->prim_count - 1;
        float scale = (float)b->bins / (cmax[axis] - cmin[axis]);
        while (i <= j) {
            const float* c = &b->cen[3 * (size_t)b->idx[i]];
            int32_t bi = bin_of(c[axis], cmin[axis], scale, b->bins);
            if ((uint32_t)bi < split) {
                i++;
            } else {
                uint32_t t = b->idx[i]; b->idx[i] = b->idx[j]; b->idx[j] = t; j--;
            }
        }
        uint32_t nl = (uint32_t)i - node->left_first;
        if (nl == 0 || nl == node->prim_count) continue;

        uint32_t li = b->node_at++;
        uint32_t ri = b->node_at++;
        b->nodes[li].left_first = node->left_first;
        b->nodes[li].prim_count = nl;
        b->nodes[ri].left_first = (uint32_t)i;
        b->nodes[ri].prim_count = node->prim_count - nl;
        node->left_first = li;
        node->prim_count = 0;

        if (sp + 2 > cap) {
            cap *= 2;
            uint32_t* ns = (uint32_t*)realloc(stack, cap * sizeof(uint32_t));
            if (!ns) { free(stack); return ORC_ECAP; }
            stack = ns;
        }
        stack[sp++] = ri;   /* popped after the whole left subtree */
        stack[sp++] = li;
    }
    free(stack);
    return ORC_OK;
}

int orc_blas_build(const float* pos9, uint32_t n, uint32_t bins, int single_prim,
                   orc_node* nodes_out, uint32_t* idx_out, uint32_t* node_count)
{
    if (!pos9 || !nodes_out || !idx_out || !node_count || n == 0 || bins < 2) return ORC_EINVAL;
    blas_state b;
    memset(&b, 0, sizeof(b));
    b.tris = (const orc_tri*)pos9;
    b.idx = idx_out;
    b.nodes = nodes_out;
    b.bins = bins;
    b.single_prim = single_prim;
    b.cen = (float*)malloc(sizeof(float) * 3 * (size_t)n);
    b.bin_min = (float*)malloc(sizeof(float) * 3 * bins);
    b.bin_max = (float*)malloc(sizeof(float) * 3 * bins);
    b.bin_cnt = (uint32_t*)malloc(sizeof(uint32_t) * bins);
    b.left_area = (float*)malloc(sizeof(float) * bins);
    b.right_area = (float*)malloc(sizeof(float) * bins);
    if (!b.cen || !b.bin_min || !b.bin_max || !b.bin_cnt || !b.left_area || !b.right_area) {
        free(b.cen); free(b.bin_min); free(b.bin_max); free(b.bin_cnt); free(b.left_area); free(b.right_area);
        return ORC_ECAP;
    }
    memset(nodes_out, 0, sizeof(orc_node) * 2 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) idx_out[i] = i;
    for (uint32_t i = 0; i < n; ++i) {
        const orc_tri* t = &b.tris[i];
        /* get_triangle_centroid, bvh_builder.cpp:252-255: (p0+p1+p2)*0.3333f */
        for (int k = 0; k < 3; ++k) b.cen[3 * (size_t)i + k] = ((t->p0[k] + t->p1[k]) + t->p2[k]) * 0.3333f;
    }
    b.nodes[0].left_first = 0;
    b.nodes[0].prim_count = n;
    b.node_at = 2;   /* node 1 is skipped (bvh_builder.cpp:44-45) */
    int rc = subdivide_all(&b, 0);
    *node_count = b.node_at;
    free(b.cen); free(b.bin_min); free(b.bin_max); free(b.bin_cnt); free(b.left_area); free(b.right_area);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* TLAS builder: tlas_builder.cpp:6-68, 86-110                              */
/* ------------------------------------------------------------------------ */

static uint32_t find_best_match(const orc_tnode* nodes, uint32_t A, const uint32_t* ids, uint32_t count, uint32_t metric)
{
    float smallest = FLT_MAX;
    uint32_t best = ~0u;
    for (uint32_t B = 0; B < count; ++B) {
        if (B == A) continue;
        const orc_tnode* na = &nodes[ids[A]];
        const orc_tnode* nb = &nodes[ids[B]];
        float area;
        if (metric == ORC_TLAS_UNION) {
            /* half area of the union box (Walter et al. 2008) */
            float mn[3], mx[3];
            for (int k = 0; k < 3; ++k) {
                mn[k] = gmin(na->aabb_min[k], nb->aabb_min[k]);
                mx[k] = gmax(na->aabb_max[k], nb->aabb_max[k]);
            }
            area = half_area(mn, mx);
        } else {
            /* the reference's metric: max of the mins and max of the maxes */
            float bmax[3], bmin[3];
            for (int k = 0; k < 3; ++k) {
                bmax[k] = gmax(na->aabb_min[k], nb->aabb_min[k]);
                bmin[k] = gmax(na->aabb_max[k], nb->aabb_max[k]);
            }
            area = half_area(bmin, bmax);
        }
        if (area < smallest) { smallest = area; best = B; }
    }
    return best;
}

int orc_tlas_build(const orc_instance* inst, uint32_t m, orc_tnode* nodes, uint32_t* node_count)
{
    return orc_tlas_build_metric(inst, m, ORC_TLAS_REFERENCE, nodes, node_count);
}

/* The same walk with either metric.  The reference's walk keeps A's index
 * when A was the last list entry and moved into B's slot (its list then
 * holds a stale A past the end); the union metric follows the moved entry. */
int orc_tlas_build_metric(const orc_instance* inst, uint32_t m, uint32_t metric, orc_tnode* nodes, uint32_t* node_count)
{
    if (!inst || !nodes || !node_count || m == 0 || metric > ORC_TLAS_UNION) return ORC_EINVAL;
    if (m > 32767) return ORC_ECAP;   /* 16-bit packed child indices */
    memset(nodes, 0, sizeof(orc_tnode) * 2 * (size_t)m);
    uint32_t* ids = (uint32_t*)calloc(m, sizeof(uint32_t));
    if (!ids) return ORC_ECAP;
    uint32_t count = m;
    uint32_t at = 1;
    for (uint32_t i = 0; i < m; ++i) {
        ids[i] = at;
        memcpy(nodes[at].aabb_min, inst[i].aabb_min, 12);
        memcpy(nodes[at].aabb_max, inst[i].aabb_max, 12);
        nodes[at].instance_idx = i;
        nodes[at].left_right = 0;
        at++;
    }
    uint32_t A = 0;
    uint32_t B = find_best_match(nodes, A, ids, count, metric);
    uint64_t guard = 0, guard_max = (uint64_t)m * m * 4 + 16;
    while (count > 1) {
        if (B == ~0u || ++guard > guard_max) { free(ids); return ORC_EINVAL; }
        uint32_t C = find_best_match(nodes, B, ids, count, metric);
        if (A == C) {
            uint32_t ia = ids[A], ib = ids[B];
            orc_tnode* nn = &nodes[at];
            nn->left_right = ia + (ib << 16);
            for (int k = 0; k < 3; ++k) {
                nn->aabb_min[k] = gmin(nodes[ia].aabb_min[k], nodes[ib].aabb_min[k]);
                nn->aabb_max[k] = gmax(nodes[ia].aabb_max[k], nodes[ib].aabb_max[k]);
            }
            ids[A] = at++;
            ids[B] = ids[count - 1];
            if (metric == ORC_TLAS_UNION && A == count - 1) A = B;
            B = find_best_match(nodes, A, ids, --count, metric);
        } else {
            A = B;
            B = C;
        }
    }
    nodes[0] = nodes[ids[A]];
    *node_count = at;
    free(ids);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* glm helpers (column-major, m[c*4+r])                                      */
/* ------------------------------------------------------------------------ */

/* mat4 * vec4, extern/glm/detail/type_mat4x4.inl:537-573: (c0*x + c1*y) + (c2*z + c3*w) */
void orc_mat4_mul_vec4(const float m[16], const float v[4], float out[4])
{
    for (int r = 0; r < 4; ++r) {
        float a0 = m[0 * 4 + r] * v[0] + m[1 * 4 + r] * v[1];
        float a1 = m[2 * 4 + r] * v[2] + m[3 * 4 + r] * v[3];
        out[r] = a0 + a1;
    }
}

void orc_mat4_mul(const float a[16], const float b[16], float out[16])
{
    float tmp[16];
    for (int c = 0; c < 4; ++c) orc_mat4_mul_vec4(a, &b[c * 4], &tmp[c * 4]);
    memcpy(out, tmp, sizeof(tmp));
}

/* compute_inverse<4,4>, extern/glm/detail/func_matrix.inl:388-446 */
void orc_mat4_inverse(const float mm[16], float out[16])
{
#define M(c, r) mm[(c) * 4 + (r)]
    float c00 = M(2,2) * M(3,3) - M(3,2) * M(2,3);
    float c02 = M(1,2) * M(3,3) - M(3,2) * M(1,3);
    float c03 = M(1,2) * M(2,3) - M(2,2) * M(1,3);
    float c04 = M(2,1) * M(3,3) - M(3,1) * M(2,3);
    float c06 = M(1,1) * M(3,3) - M(3,1) * M(1,3);
    float c07 = M(1,1) * M(2,3) - M(2,1) * M(1,3);
    float c08 = M(2,1) * M(3,2) - M(3,1) * M(2,2);
    float c10 = M(1,1) * M(3,2) - M(3,1) * M(1,2);
    float c11 = M(1,1) * M(2,2) - M(2,1) * M(1,2);
    float c12 = M(2,0) * M(3,3) - M(3,0) * M(2,3);
    float c14 = M(1,0) * M(3,3) - M(3,0) * M(1,3);
    float c15 = M(1,0) * M(2,3) - M(2,0) * M(1,3);
    float c16 = M(2,0) * M(3,2) - M(3,0) * M(2,2);
    float c18 = M(1,0) * M(3,2) - M(3,0) * M(1,2);
    float c19 = M(1,0) * M(2,2) - M(2,0) * M(1,2);
    float c20 = M(2,0) * M(3,1) - M(3,0) * M(2,1);
    float c22 = M(1,0) * M(3,1) - M(3,0) * M(1,1);
    float c23 = M(1,0) * M(2,1) - M(2,0) * M(1,1);
    float f0[4] = { c00, c00, c02, c03 }, f1[4] = { c04, c04, c06, c07 }, f2[4] = { c08, c08, c10, c11 };
    float f3[4] = { c12, c12, c14, c15 }, f4[4] = { c16, c16, c18, c19 }, f5[4] = { c20, c20, c22, c23 };
    float v0[4] = { M(1,0), M(0,0), M(0,0), M(0,0) };
    float v1[4] = { M(1,1), M(0,1), M(0,1), M(0,1) };
    float v2[4] = { M(1,2), M(0,2), M(0,2), M(0,2) };
    float v3[4] = { M(1,3), M(0,3), M(0,3), M(0,3) };
    float inv[16];
    const float sa[4] = { 1, -1, 1, -1 }, sb[4] = { -1, 1, -1, 1 };
    for (int k = 0; k < 4; ++k) {
        inv[0 * 4 + k] = ((v1[k] * f0[k] - v2[k] * f1[k]) + v3[k] * f2[k]) * sa[k];
        inv[1 * 4 + k] = ((v0[k] * f0[k] - v2[k] * f3[k]) + v3[k] * f4[k]) * sb[k];
        inv[2 * 4 + k] = ((v0[k] * f1[k] - v1[k] * f3[k]) + v3[k] * f5[k]) * sa[k];
        inv[3 * 4 + k] = ((v0[k] * f2[k] - v1[k] * f4[k]) + v2[k] * f5[k]) * sb[k];
    }
    float row0[4] = { inv[0], inv[4], inv[8], inv[12] };
    float d0[4];
    for (int k = 0; k < 4; ++k) d0[k] = M(0, k) * row0[k];
    float det = (d0[0] + d0[1]) + (d0[2] + d0[3]);
    float one_over = 1.0f / det;
    for (int i = 0; i < 16; ++i) out[i] = inv[i] * one_over;
#undef M
}

/* submit_render_mesh software branch, renderer.cpp:1318-1336.  The record's
 * bounds start zeroed (begin_frame allocates it with ARENA_ALLOC_ARRAY_ZERO,
 * renderer.cpp:563-566), so the world AABB always contains the origin. */
void orc_instance_record(const float transform[16], const float bmin[3], const float bmax[3],
                         uint32_t bvh_index, orc_instance* out)
{
    memset(out, 0, sizeof(*out));
    orc_mat4_inverse(transform, out->world_to_local);
    out->bvh_index = bvh_index;
    for (uint32_t i = 0; i < 8; ++i) {
        float p[4] = { (i & 1) ? bmax[0] : bmin[0], (i & 2) ? bmax[1] : bmin[1], (i & 4) ? bmax[2] : bmin[2], 1.0f };
        float w[4];
        orc_mat4_mul_vec4(transform, p, w);
        for (int k = 0; k < 3; ++k) {
            out->aabb_min[k] = gmin(out->aabb_min[k], w[k]);
            out->aabb_max[k] = gmax(out->aabb_max[k], w[k]);
        }
    }
}
