/*
 * CPU oracle for the wavefront path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * algorithms, used by tests/, by the smoke check and by bench.py's CPU
 * baseline leg as the thing results are checked against.  Nothing in the
 * product library links, loads or calls it.
 *
 * Reference files restated (paths relative to the reference root):
 *   source/renderer/bvh/bvh_builder.cpp, bvh/as_util.h     -> orc_blas_build
 *   source/renderer/bvh/tlas_builder.cpp                   -> orc_tlas_build
 *   source/renderer/renderer.cpp:1318-1336                 -> orc_instance_record
 *   source/renderer/shaders/intersect.hlsl:7-76            -> orc_intersect_*
 *   source/renderer/shaders/accelstruct.hlsl:58-218        -> orc_trace_closest
 *   source/renderer/shaders/common.hlsl:145-169            -> orc_primary_rays
 *   extern/glm (v1.0.2) lookAtLH / perspectiveFovLH_ZO /
 *     inverse / mat4*vec4                                   -> orc_mat4_*, orc_camera_*
 *   source/renderer/shaders/wavefront/{generate,extend,shade}.hlsl,
 *     renderer.cpp:742-939                                  -> orc_render (wavefront loop)
 *
 * Floating point: FP32, no contraction (built with -ffp-contract=off), no
 * fast-math, SSE scalar arithmetic, correctly rounded divide and sqrt.
 *
 * Parity status per function is listed in the pinning table of the test
 * suite (tests/test_oracle_*.py); functions without an independent check say
 * "parity unpinned" in their comment below.
 */
#ifndef WF_ORACLE_H
#define WF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reference wire layouts (shared.hlsl.h:106-161), 4-byte packed. */
typedef struct { float aabb_min[3]; uint32_t left_first; float aabb_max[3]; uint32_t prim_count; } orc_node;      /* bvh_node_t 32 B */
typedef struct { float p0[3]; float p1[3]; float p2[3]; } orc_tri;                                                  /* bvh_triangle_t 36 B */
typedef struct { float aabb_min[3]; uint32_t left_right; float aabb_max[3]; uint32_t instance_idx; } orc_tnode;   /* tlas_node_t 32 B */
typedef struct { float world_to_local[16]; float aabb_min[3]; float aabb_max[3]; uint32_t bvh_index; } orc_instance; /* bvh_instance_t 92 B */
typedef struct { uint32_t instance_idx; uint32_t primitive_idx; float t; float bary[2]; } orc_hit;                 /* hit_result_t 20 B */
typedef struct { float origin[3]; float tmin; float dir[3]; float tmax; } orc_ray;                                 /* RayDesc2 32 B */
typedef struct { uint32_t tlas_nodes, blas_nodes, instances, tri_tests; } orc_counters;

enum { ORC_OK = 0, ORC_EINVAL = 1, ORC_ECAP = 2, ORC_ESTACK = 5 };

/* ---- builders ---------------------------------------------------------- */
/* SAH binned BLAS.  pos9 = N triangles x (p0,p1,p2).  nodes_out holds 2N
 * entries, idx_out N entries.  *node_count = nodes written (node_at). */
int orc_blas_build(const float* pos9, uint32_t n, uint32_t bins, int single_prim,
                   orc_node* nodes_out, uint32_t* idx_out, uint32_t* node_count);

/* Agglomerative TLAS over M instance records; nodes_out holds 2M entries. */
int orc_tlas_build(const orc_instance* inst, uint32_t m, orc_tnode* nodes_out, uint32_t* node_count);
/* The same agglomerative walk with a choice of match metric: the reference's
 * (max of mins / max of maxes, as orc_tlas_build) or the half area of the
 * union box.  The union metric has no counterpart in the reference. */
enum { ORC_TLAS_REFERENCE = 0, ORC_TLAS_UNION = 1 };
int orc_tlas_build_metric(const orc_instance* inst, uint32_t m, uint32_t metric, orc_tnode* nodes_out, uint32_t* node_count);

/* glm restatements (column-major m[col*4+row]) */
void orc_mat4_inverse(const float m[16], float out[16]);
void orc_mat4_mul_vec4(const float m[16], const float v[4], float out[4]);
void orc_mat4_mul(const float a[16], const float b[16], float out[16]);

/* Software instance record from a local_to_world transform and the BLAS
 * root bounds (renderer.cpp:1318-1336). */
void orc_instance_record(const float transform[16], const float blas_min[3], const float blas_max[3],
                         uint32_t bvh_index, orc_instance* out);

/* ---- intersection primitives ------------------------------------------- */
int   orc_intersect_triangle(const float v0[3], const float v1[3], const float v2[3],
                             const float o[3], const float d[3], float* t_inout, float bary[2]);
float orc_intersect_aabb(const float bmin[3], const float bmax[3], const float o[3],
                         const float inv_dir[3], float t);

/* ---- scene + traversal -------------------------------------------------- */
typedef struct {
    const orc_tnode*    tnodes;     uint32_t n_tnodes;
    const orc_instance* inst;       uint32_t n_inst;
    const orc_node*     nodes;      const uint32_t* node_off;   /* per BLAS */
    const orc_tri*      tris;       const uint32_t* idx;        const uint32_t* tri_off; /* per BLAS */
    uint32_t            n_blas;
} orc_scene;

/* Closest hit for each ray (extend.hlsl:25-40).  counters may be NULL
 * (else one orc_counters per ray).  nthreads <= 1 runs on the caller. */
int orc_trace_closest(const orc_scene* s, const orc_ray* rays, uint32_t n, orc_hit* hits,
                      orc_counters* counters, int nthreads);

/* Occlusion (any hit with t in [0, tmax)); visible[i] = 1 when unoccluded. */
int orc_trace_any(const orc_scene* s, const orc_ray* rays, uint32_t n, uint32_t* visible, int nthreads);

/* Brute force closest t over every instance and triangle (no BVH). */
int orc_brute_closest(const orc_scene* s, const orc_ray* rays, uint32_t n, orc_hit* hits);

/* ---- camera + generate --------------------------------------------------- */
typedef struct {
    float world_to_view[16]; float view_to_world[16]; float view_to_clip[16]; float clip_to_view[16];
    float render_dim[2]; float near_plane; float far_plane;
} orc_view;  /* view_t 272 B */

void orc_camera_view(const float eye[3], const float target[3], float vfov_deg,
                     uint32_t width, uint32_t height, orc_view* out);
void orc_primary_rays(const orc_view* v, uint32_t width, uint32_t height, orc_ray* rays);

/* ---- wavefront render (generate/extend/shade/connect) -------------------- */
/* Per-instance shading record (instance_data_t, shared.hlsl.h:87-95). */
typedef struct {
    float    local_to_world[16];
    float    base_color[3];
    float    emissive[3];          /* emissive_factor * emissive_strength */
    int32_t  base_color_tex;       /* -1: none (white) */
    uint32_t mesh;
    int32_t  emissive_tex;         /* -1: none; emission = emissive * texel (sample_material, material.hlsl:35-37) */
    int32_t  normal_tex;           /* -1: none; read by the debug view modes only */
    int32_t  metal_rough_tex;      /* -1: none; read by the debug view modes only */
} orc_material_instance;

typedef struct {
    const float*    shade_tris;    /* all meshes: 24 floats per triangle: p0 p1 p2 (9), n0 n1 n2 (9), uv0 uv1 uv2 (6) */
    const uint32_t* shade_off;     /* per mesh, first triangle */
    const uint32_t* tri_count;     /* per mesh */
    const orc_material_instance* mi;   /* per instance */
    uint32_t        n_inst;
    const float*    tex_rgba;      /* RGBA float textures, concatenated */
    const uint32_t* tex_off;       /* per texture, offset in floats */
    const uint32_t* tex_w;
    const uint32_t* tex_h;
    float           env[3];        /* constant environment radiance (hdr_env_strength x white) */
    int32_t         env_tex;       /* equirect environment texture (sample.hlsl:4-11), -1: the constant env */
    float           env_strength;  /* hdr_env_strength, with env_tex */
} orc_shading;

/* Texel index of the environment texture in direction d: uv = (atan2(z, x),
 * asin(-y)) * (0.1591, 0.3183) + 0.5 (direction_to_equirect_uv,
 * sample.hlsl:4-11), point load at uint2(uv * size) (shade.hlsl:36-40). */
uint32_t orc_env_texel(const orc_shading* sh, const float d[3]);

typedef struct {
    uint32_t max_bounces;
    uint32_t spp;
    uint32_t rr_start;             /* bounce index at which Russian roulette starts */
    uint32_t nee;                  /* 1: next-event estimation through connect */
    uint32_t frame_seed;
    float    rr_survival;
    uint32_t path_base;            /* id of the first path (scenes of a batch number paths consecutively) */
    uint32_t uniform_diffuse;      /* 1: uniform hemisphere sampling (cosine_weighted_diffuse = 0), else cosine */
    uint32_t view_mode;            /* RENDER_VIEW_MODE (shared.hlsl.h:17-32), 0 = the path tracer */
} orc_render_params;

/* energy_rgba: width*height*4 floats (the resolved per-pixel mean over spp). */
int orc_render(const orc_scene* s, const orc_shading* sh, const orc_view* v, uint32_t width, uint32_t height,
               const orc_render_params* p, float* energy_rgba, int nthreads);

/* The emitter table next-event estimation samples: per emissive triangle
 * (instance, primitive) and the cumulative area fraction; returns the count
 * (call with NULL outputs for the count) and the total world area. */
uint32_t orc_emitters(const orc_shading* sh, uint32_t* ids, float* cdf, float* total_area);

/* ---- differentiable rendering (FP64 reference) -------------------------------
 * The FP32 path walker of orc_render records, per path vertex, everything
 * that does not depend on albedo or emission (paths are sampled
 * independently of both: cosine sampling, constant Russian roulette).  The
 * FP64 evaluator re-weights the recorded paths for any texel values and
 * emission strengths; the FP64 adjoint is path-replay backpropagation over
 * the same records. */
enum { ORC_V_END = 0, ORC_V_MISS = 1, ORC_V_EMIT = 2, ORC_V_DIFFUSE = 3 };
typedef struct {
    uint32_t kind;          /* ORC_V_* */
    uint32_t inst;          /* hit instance */
    uint32_t emit_counted;  /* ORC_V_EMIT: emission added (no NEE, or camera ray) */
    uint32_t rr;            /* throughput divided by rr_survival after this vertex */
    float    noL, inv_pdf;  /* bounce weight (NoL * (a / pi)) * inv_pdf */
    int32_t  tex;           /* base-color texture of the instance, -1 none */
    uint32_t texel[4];      /* bilinear corners (x0,y0) (x1,y0) (x0,y1) (x1,y1), texel index in the texture */
    float    fx, fy;        /* bilinear fractions */
    uint32_t nee;           /* 1: a visible shadow-ray connection to light_inst */
    uint32_t light_inst;
    float    nee_wgt;       /* G * total emitter area */
    float    em_tex[3];     /* ORC_V_EMIT: emissive texture value at the hit (1 without a texture) */
    float    nee_tex[3];    /* nee: emissive texture value at the light sample (1 without a texture) */
} orc_vertex;

/* Walk every path once (FP32 geometry, as orc_render); rec holds
 * (max_bounces + 1) vertices per path, terminated by ORC_V_END. */
int orc_record_paths(const orc_scene* s, const orc_shading* sh, const orc_view* v, uint32_t width, uint32_t height,
                     const orc_render_params* p, orc_vertex* rec, int nthreads);

/* FP64 re-weighting of recorded paths.  texels: RGBA doubles of texture
 * `tex` (other textures use their float texels); strength: per instance,
 * the emissive_strength (emission = emissive_factor * strength).
 * image: width*height*3 doubles (mean over spp). */
int orc_eval_f64(const orc_shading* sh, const orc_render_params* p, uint32_t width, uint32_t height,
                 const orc_vertex* rec, int32_t tex, const double* texels, const double* factor3,
                 const double* strength, double* image);

/* FP64 path-replay adjoint: given dLoss/dimage (width*height*3), the
 * gradients with respect to the texels of `tex` (RGBA doubles) and the
 * per-instance emission strengths.  Both functions return ORC_EINVAL when
 * `tex` is also an emissive or the environment texture. */
int orc_adjoint_f64(const orc_shading* sh, const orc_render_params* p, uint32_t width, uint32_t height,
                    const orc_vertex* rec, int32_t tex, const double* texels, const double* factor3,
                    const double* strength, const double* dl_dimage, double* d_texels, double* d_strength);

/* ---- post-process (post_process.hlsl:25-72) ---------------------------------
 * One frame over n pixels: energy_rgba (the frame's per-pixel mean) and the
 * accumulator (in/out) -> RGBA8 final colour (R in the low byte); final_rgba,
 * when not NULL, receives the final colour before the 8-bit store. */
void orc_post_process(const float* energy_rgba, uint32_t n, uint32_t sample_count, uint32_t accumulate,
                      uint32_t view_mode, float* accum_rgba, uint32_t* final_rgba8, float* final_rgba);

/* RNG and trigonometry used by shade (exposed for known-answer tests). */
uint32_t orc_pcg32_first(uint64_t initstate, uint64_t initseq);
void     orc_path_rng(uint32_t path, uint32_t bounce, uint32_t frame_seed, float out[6]);
float    orc_sin_turns(float u);
float    orc_cos_turns(float u);
void     orc_cosine_sample(const float n[3], float r0, float r1, float out[3]);
void     orc_uniform_sample(const float n[3], float r0, float r1, float out[3]);
uint32_t orc_murmur_mix(uint32_t h);

#ifdef __cplusplus
}
#endif
#endif
