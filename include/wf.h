/*
 * wf.h — C-ABI of the MI355X wavefront path tracer.
 *
 * The entry points mirror the reference renderer's front-end
 * (source/renderer/renderer.h:13-55) so that a caller of that API can switch
 * to this library; the global singleton `g_renderer` (renderer.cpp:25)
 * becomes an explicit wf_renderer handle.  Every entry point returns a
 * wf_status instead of asserting (core/assertion.h:51-67), copies its inputs
 * before returning, and never retains caller pointers.  Output buffers are
 * caller-owned host memory unless the name ends in _device.
 *
 * Wire formats are the reference's (shared.hlsl.h:58-161): wf_vertex is
 * vertex_t (48 B), wf_ray is RayDesc2 (32 B), wf_hit is hit_result_t (20 B),
 * wf_bvh_node is bvh_node_t (32 B), wf_tlas_node is tlas_node_t (32 B),
 * wf_bvh_instance is bvh_instance_t (92 B).  All structs are 4-byte packed.
 *
 * Threading: one wf_renderer per thread; a renderer is not internally
 * synchronised.  All GPU work of a renderer is ordered on its own HIP stream;
 * readbacks synchronise that stream.
 */
#ifndef WF_H
#define WF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WF_API __attribute__((visibility("default")))

typedef enum wf_status {
    WF_OK = 0,
    WF_ERROR_INVALID_ARGUMENT = 1,
    WF_ERROR_CAPACITY = 2,          /* an instance/texture/mesh/TLAS limit was exceeded */
    WF_ERROR_OUT_OF_MEMORY = 3,
    WF_ERROR_DEVICE = 4,            /* a HIP runtime call failed */
    WF_ERROR_STACK_OVERFLOW = 5,    /* a traversal stack overflowed (result for that ray is a miss) */
    WF_ERROR_INVALID_STATE = 6,     /* call order violated (e.g. render without begin_scene) */
    WF_ERROR_NOT_FOUND = 7          /* stale or unknown handle */
} wf_status;

typedef struct wf_renderer wf_renderer;

/* resource_handle_t (core/api_types.h:4-25): low 32 bits = index, high 32 bits
 * = version; WF_INVALID_HANDLE = UINT64_MAX. */
typedef uint64_t wf_mesh_handle;
typedef uint64_t wf_texture_handle;
#define WF_INVALID_HANDLE UINT64_MAX

/* vertex_t (shared.hlsl.h:58-64) */
typedef struct wf_vertex { float position[3]; float normal[3]; float tangent[4]; float uv[2]; } wf_vertex;
/* RayDesc2 (shared.hlsl.h:155-161).  The software path ignores tmin (the
 * origin is already nudged by direction*1e-8, common.hlsl:145-154) and starts
 * the closest-hit search with ray.t = tmax. */
typedef struct wf_ray { float origin[3]; float tmin; float direction[3]; float tmax; } wf_ray;
/* hit_result_t (shared.hlsl.h:144-150); a miss is {~0u, ~0u, FLT_MAX, {0,0}} */
typedef struct wf_hit { uint32_t instance_idx; uint32_t primitive_idx; float t; float bary[2]; } wf_hit;
/* bvh_header_t / bvh_node_t / bvh_triangle_t / bvh_instance_t / tlas_header_t / tlas_node_t */
typedef struct wf_bvh_header { uint32_t nodes_offset, triangles_offset, indices_offset; } wf_bvh_header;
typedef struct wf_bvh_node { float aabb_min[3]; uint32_t left_first; float aabb_max[3]; uint32_t prim_count; } wf_bvh_node;
typedef struct wf_bvh_triangle { float p0[3], p1[3], p2[3]; } wf_bvh_triangle;
typedef struct wf_bvh_instance { float world_to_local[16]; float aabb_min[3]; float aabb_max[3]; uint32_t bvh_index; } wf_bvh_instance;
typedef struct wf_tlas_header { uint32_t nodes_offset, instances_offset; } wf_tlas_header;
typedef struct wf_tlas_node { float aabb_min[3]; uint32_t left_right; float aabb_max[3]; uint32_t instance_idx; } wf_tlas_node;

/* renderer::init_params_t (renderer.h:13-19) */
typedef struct wf_init_params { uint32_t render_width; uint32_t render_height; uint32_t backbuffer_count; uint32_t vsync; } wf_init_params;

/* camera_t (core/camera/camera.h:5-13): column-major glm matrices */
typedef struct wf_camera { float inv_view_matrix[16]; float view_matrix[16]; float vfov_deg; float yaw; float pitch; } wf_camera;

/* view_t (shared.hlsl.h:46-56): the per-frame camera constants handed to the kernels */
typedef struct wf_view {
    float world_to_view[16]; float view_to_world[16]; float view_to_clip[16]; float clip_to_view[16];
    float render_dim[2]; float near_plane; float far_plane;
} wf_view;

/* TEXTURE_FORMAT subset accepted by create_render_texture (renderer_fwd.h:16-34;
 * the block-compressed BC formats are not accepted) */
enum { WF_TEXTURE_FORMAT_RGBA8 = 1, WF_TEXTURE_FORMAT_RGBA8_SRGB = 2, WF_TEXTURE_FORMAT_RGB32F = 3, WF_TEXTURE_FORMAT_RGBA32F = 4,
       WF_TEXTURE_FORMAT_RG8 = 5 };

/* renderer::render_texture_params_t (renderer.h:33-42) */
typedef struct wf_render_texture_params { uint32_t width; uint32_t height; uint32_t bits_per_pixel; uint32_t format; const void* ptr_data; } wf_render_texture_params;

/* renderer::render_mesh_params_t (renderer.h:45-53) */
typedef struct wf_render_mesh_params { uint32_t vertex_count; const wf_vertex* vertices; uint32_t index_count; const uint32_t* indices; } wf_render_mesh_params;

/* material_asset_t (core/assets/asset_types.h:10-21) */
typedef struct wf_material_asset {
    float base_color_factor[4]; float metallic_factor; float roughness_factor;
    float emissive_factor[3]; float emissive_strength;
    wf_texture_handle base_color_texture; wf_texture_handle normal_texture;
    wf_texture_handle metallic_roughness_texture; wf_texture_handle emissive_texture;
} wf_material_asset;

/* render_settings_t (shared.hlsl.h:34-44) followed by this library's additions.
 * Defaults (renderer.cpp:32-45 plus additions): wavefront on, software RT on
 * (the only traversal this library has), view mode none, max_bounces 3,
 * cosine-weighted diffuse, accumulate, env strength 1, spp 1, Russian
 * roulette from bounce 3 with survival 0.75, next-event estimation off,
 * frame_seed 0 (0 = advance an internal xorshift per frame, renderer.cpp:739). */
/* RENDER_VIEW_MODE (shared.hlsl.h:17-32): debug views that replace a camera
 * ray's radiance with one property of what it hits (shade.hlsl:140-153). */
typedef enum wf_render_view_mode {
    WF_RENDER_VIEW_MODE_NONE = 0,
    WF_RENDER_VIEW_MODE_GEOMETRY_INSTANCE, WF_RENDER_VIEW_MODE_GEOMETRY_PRIMITIVE, WF_RENDER_VIEW_MODE_GEOMETRY_BARYCENTRICS,
    WF_RENDER_VIEW_MODE_GEOMETRY_NORMAL, WF_RENDER_VIEW_MODE_GEOMETRY_UV, WF_RENDER_VIEW_MODE_MATERIAL_BASE_COLOR,
    WF_RENDER_VIEW_MODE_MATERIAL_NORMAL, WF_RENDER_VIEW_MODE_MATERIAL_METALLIC_ROUGHNESS, WF_RENDER_VIEW_MODE_MATERIAL_EMISSIVE,
    WF_RENDER_VIEW_MODE_WORLD_NORMAL, WF_RENDER_VIEW_MODE_RENDER_TARGET_DEPTH,
    WF_RENDER_VIEW_MODE_COUNT
} wf_render_view_mode;

typedef struct wf_render_settings {
    uint32_t use_wavefront_pathtracing;
    uint32_t use_software_rt;
    uint32_t render_view_mode;          /* wf_render_view_mode */
    uint32_t max_bounces;
    uint32_t cosine_weighted_diffuse;   /* 1: cosine-weighted hemisphere sampling; 0: uniform (shade.hlsl:114-131) */
    uint32_t accumulate;
    float    hdr_env_strength;
    /* additions (no reference counterpart) */
    uint32_t spp;
    uint32_t rr_start_bounce;
    float    rr_survival;
    uint32_t next_event_estimation;
    uint32_t frame_seed;
    uint32_t use_hip_graph;      /* 1: the frame's launches replay as one captured hipGraph */
    uint32_t separate_connect;   /* 1: shadow rays in their own launch after each shade (the reference's
                                    one launch per stage); 0: traced in the next bounce's extend launch.
                                    Same image either way. */
} wf_render_settings;

/* bvh_builder_t::build_options_t (bvh/bvh_builder.h:17-21); the reference
 * renderer uses interval_count 8, subdivide_single_prim false (renderer.cpp:95-96) */
typedef struct wf_build_options { uint32_t interval_count; uint32_t subdivide_single_prim; } wf_build_options;

/* Per-frame stage timings (ms) from HIP events, the reference's GPU profiler
 * scopes (gpu_profiler.h:13-28) for the wavefront stages. */
typedef struct wf_frame_stats {
    float ms_clear, ms_generate, ms_extend, ms_shade, ms_connect, ms_post, ms_total;
    uint32_t rays_extended;      /* closest-hit rays traced this frame */
    uint32_t shadow_rays;        /* occlusion rays traced this frame */
    uint32_t stack_overflows;
    /* per bounce (wavefront frames; the megakernel reports totals in [0]) */
    uint32_t rays_per_bounce[9];
    uint32_t shadow_rays_per_bounce[9];
} wf_frame_stats;

/* ---- lifecycle: renderer::init / exit (renderer.cpp:222-541) ---------- */
WF_API wf_status wf_renderer_init(const wf_init_params* params, wf_renderer** out_renderer);
WF_API wf_status wf_renderer_exit(wf_renderer* r);
WF_API const char* wf_renderer_last_error(const wf_renderer* r);
/* The hipStream_t all of the renderer's GPU work is ordered on (for callers
 * that record events or enqueue work around it). */
WF_API wf_status wf_renderer_get_stream(const wf_renderer* r, void** out_stream);

/* ---- frame / scene: renderer::begin_frame .. end_scene (renderer.cpp:543-1055) */
WF_API wf_status wf_renderer_begin_frame(wf_renderer* r);
WF_API wf_status wf_renderer_end_frame(wf_renderer* r);
WF_API wf_status wf_renderer_begin_scene(wf_renderer* r, const wf_camera* camera, wf_texture_handle env_texture);
WF_API wf_status wf_renderer_render(wf_renderer* r);
WF_API wf_status wf_renderer_end_scene(wf_renderer* r);

/* ---- batches (no reference counterpart): many scenes, each with its own
 * camera and resolution, rendered by one set of launches over shared queues.
 * begin_batch replaces begin_scene; batch_add_scene opens scene slot k
 * (returned) and submit_render_mesh adds instances to the open slot;
 * end_scene closes the batch.  render() draws every slot; render_scenes
 * draws the slots [first, first+count) only (the per-scene loop).  A scene
 * renders the same image alone or inside a batch (paths are numbered over
 * all slots). */
WF_API wf_status wf_renderer_begin_batch(wf_renderer* r);
WF_API wf_status wf_renderer_batch_add_scene(wf_renderer* r, const wf_camera* camera, uint32_t width, uint32_t height, uint32_t* out_scene);
WF_API wf_status wf_renderer_get_scene_count(const wf_renderer* r, uint32_t* out_count);
WF_API wf_status wf_renderer_render_scenes(wf_renderer* r, uint32_t first, uint32_t count);
/* linear energy of one scene slot: width*height RGBA float */
WF_API wf_status wf_renderer_read_scene_energy(wf_renderer* r, uint32_t scene, float* out_rgba);

/* ---- differentiable rendering (no reference counterpart) -----------------
 * Path-replay adjoint (Vicini, Speierer, Jakob, SIGGRAPH 2021) of the last
 * render(): every path of that frame is replayed with the same random
 * numbers and the gradients of a loss are accumulated, given dLoss/dimage
 * (RGBA floats per pixel of the frame's image, all scenes of a batch
 * concatenated; alpha ignored):
 *   grad_texels   dLoss/d(texel) of base-color texture `texture` (RGBA floats,
 *                 width*height*4; pass WF_INVALID_HANDLE to skip), summed
 *                 over every instance and scene that samples it; a texture
 *                 that is also an emissive or the environment texture is
 *                 rejected (WF_ERROR_INVALID_ARGUMENT);
 *   grad_emission dLoss/d(emissive_strength) per instance (all scenes'
 *                 instances in submission order).
 * Russian roulette and sampling do not depend on albedo or emission, so the
 * replay follows the forward paths exactly.  Requires that render() ran last
 * with the same scene, settings, texture data and emission strengths. */
WF_API wf_status wf_renderer_render_adjoint(wf_renderer* r, const float* dl_dimage, wf_texture_handle texture,
                                           float* out_grad_texels, float* out_grad_emission);
WF_API wf_status wf_renderer_render_adjoint_device(wf_renderer* r, const float* d_dl_dimage, wf_texture_handle texture,
                                                  float* d_grad_texels, float* d_grad_emission);
/* in-place texture update (RGBA floats, same size) and emission strength of
 * one instance (index over all scenes' instances), for optimisation loops */
WF_API wf_status wf_renderer_update_texture(wf_renderer* r, wf_texture_handle texture, const float* rgba);
WF_API wf_status wf_renderer_update_texture_device(wf_renderer* r, wf_texture_handle texture, const float* d_rgba);
WF_API wf_status wf_renderer_set_emission_strength(wf_renderer* r, uint32_t instance, float strength);
/* device-to-device copy of the whole frame image (all scenes' pixels) */
WF_API wf_status wf_renderer_read_energy_device(wf_renderer* r, float* d_out_rgba);

/* ---- resources: create_render_texture / create_render_mesh / submit_render_mesh
 *      (renderer.cpp:1150-1350) ---------------------------------------- */
WF_API wf_status wf_renderer_create_render_texture(wf_renderer* r, const wf_render_texture_params* params, wf_texture_handle* out_handle);
WF_API wf_status wf_renderer_create_render_mesh(wf_renderer* r, const wf_render_mesh_params* params, wf_mesh_handle* out_handle);
WF_API wf_status wf_renderer_submit_render_mesh(wf_renderer* r, wf_mesh_handle mesh, const float transform[16], const wf_material_asset* material);

/* ---- settings (render_settings_t, editable through the reference's UI, renderer.cpp:1120-1148) */
WF_API wf_status wf_renderer_get_settings(const wf_renderer* r, wf_render_settings* out);
WF_API wf_status wf_renderer_set_settings(wf_renderer* r, const wf_render_settings* settings);
/* options for meshes created after the call (renderer.cpp:93-96) */
WF_API wf_status wf_renderer_set_build_options(wf_renderer* r, const wf_build_options* options);
/* Where create_render_mesh builds a BLAS (no reference counterpart; the
 * reference builds on the CPU at load time, renderer.cpp:1266-1269).  Both
 * builders produce the same nodes and triangle order.  AUTO uses the GPU for
 * meshes of at least 16,384 triangles when interval_count is 2..64. */
typedef enum wf_blas_builder { WF_BLAS_BUILDER_AUTO = 0, WF_BLAS_BUILDER_HOST = 1, WF_BLAS_BUILDER_DEVICE = 2 } wf_blas_builder;
WF_API wf_status wf_renderer_set_blas_builder(wf_renderer* r, uint32_t builder);
/* Where the renderer builds each scene's TLAS when its instances change (the
 * reference rebuilds it on the CPU every frame, renderer.cpp:637-645); AUTO
 * uses the GPU from 4,096 instances.  Same nodes either way. */
WF_API wf_status wf_renderer_set_tlas_builder(wf_renderer* r, uint32_t builder);
/* The match metric of the renderer's TLAS builds (wf_tlas_metric; default
 * REFERENCE, the reference's tree).  Takes effect at the next TLAS build. */
WF_API wf_status wf_renderer_set_tlas_metric(wf_renderer* r, uint32_t metric);

/* Traversal kernel configuration (no reference counterpart: the autotuner's
 * axes; defaults 256, 16, 1, BVH2, no cap, 16, 0, 4, 0, 64, on).  workgroup_size 64/128/256; lds_stack_entries 0 (private stack in
 * scratch), 4, 8, 12 or 16 (LDS, spilling to global memory past that depth);
 * skip_popped_interior drops popped interior nodes whose entry distance is
 * not below the current hit distance; bvh_layout 0 = BVH2; stack_limit caps
 * the traversal stack below its 128 entries per lane (0 = no cap; a ray that
 * would exceed it reports WF_ERROR_STACK_OVERFLOW); refill_min (1..64, 0 =
 * default 16) is how many lanes of a 64-lane wave must have finished their
 * rays before the wave claims new ones; tlas_in_lds 1 copies the TLAS into
 * each workgroup's LDS when it fits (at most 1024 interior nodes and 64 KB
 * of LDS with the stack; 0 = read it through the caches);
 * interior_visits (1..8, 0 = default 4) is how many interior nodes a lane may
 * visit per pass of the traversal loop before the pass's instance-entry and
 * leaf steps.
 * big_claims selects how a wave of a closest-hit launch that also carries
 * shadow rays claims queue entries: 0 = by the queues' capacity (at least
 * 2^24 entries, as in a large batch: up to 256 entries per atomic; fewer:
 * one entry per idle lane), 1 = always the large claims, 2 = never.
 * claim_segments (a power of two in 1..64, 0 = default 64) caps how many
 * cursors an extend launch claims its queue from: 64-entry chunks of the
 * queue are dealt round-robin to that many segments (fewer for queues under
 * 512 entries per segment), each with a cursor of its own.
 * cost_order (0 = default on, 1 = on, 2 = off; get returns 1 or 2) lets the
 * trace calls on caller rays (wf_trace_closest, _occluded, _counters,
 * _device) claim the rays predicted to traverse longest first: a pre-pass
 * buckets the rays by their chord through the scene bounds, weighted by its
 * distance from the bounds' centre, so long traversal chains start early
 * instead of trailing the launch.  Results are the same either way; queues
 * of at most two rays per persistent lane are traced in the callers' order. */
typedef struct wf_traversal_config {
    uint32_t workgroup_size; uint32_t lds_stack_entries; uint32_t skip_popped_interior; uint32_t bvh_layout;
    uint32_t stack_limit; uint32_t refill_min; uint32_t tlas_in_lds; uint32_t interior_visits;
    uint32_t big_claims; uint32_t claim_segments; uint32_t cost_order;
} wf_traversal_config;
WF_API wf_status wf_renderer_set_traversal_config(wf_renderer* r, const wf_traversal_config* config);
WF_API wf_status wf_renderer_get_traversal_config(const wf_renderer* r, wf_traversal_config* out);

/* ---- camera::create (core/camera/camera.cpp:8-16): lookAtLH, up = +y ---- */
WF_API wf_status wf_camera_create(wf_camera* out, const float view_origin[3], const float target_pos[3], float vfov_deg);
/* the view_t that begin_scene derives (renderer.cpp:614-629) */
WF_API wf_status wf_renderer_get_view(const wf_renderer* r, wf_view* out);
/* the same derivation without a renderer (host only) */
WF_API wf_status wf_make_view(const wf_camera* camera, uint32_t width, uint32_t height, wf_view* out);

/* ---- readbacks --------------------------------------------------------- */
/* linear path-traced energy of the last render(): width*height RGBA float
 * (the reference's texture_energy, renderer.cpp:470-474) */
WF_API wf_status wf_renderer_read_energy(wf_renderer* r, float* out_rgba);
/* progressive accumulator (post_process.hlsl:40-51): width*height RGBA float */
WF_API wf_status wf_renderer_read_accumulator(wf_renderer* r, float* out_rgba);
/* display output (post_process.hlsl:53-71): width*height RGBA8; Reinhard
 * for view mode 0 only, the sRGB curve for modes 0, 6 and 9 only */
WF_API wf_status wf_renderer_read_final_color(wf_renderer* r, uint8_t* out_rgba8);
WF_API wf_status wf_renderer_get_frame_stats(const wf_renderer* r, wf_frame_stats* out);
/* image files (host only; no reference counterpart, which presents to a
 * swapchain): linear RGBA float -> PFM (RGB, little endian, bottom row
 * first); RGBA8 -> binary PPM (RGB) */
WF_API wf_status wf_write_pfm(const char* path, const float* rgba, uint32_t width, uint32_t height);
WF_API wf_status wf_write_ppm(const char* path, const uint8_t* rgba8, uint32_t width, uint32_t height);

/* ---- the extend stage on caller rays (extend.hlsl:25-40) -------------- */
/* Builds the TLAS from the instances submitted since begin_scene and uploads
 * it; render() does this itself (renderer.cpp:637-670). */
WF_API wf_status wf_renderer_build_tlas(wf_renderer* r);
/* World bounds of the built TLAS (its root node). */
WF_API wf_status wf_renderer_get_scene_bounds(const wf_renderer* r, float out_min[3], float out_max[3]);
/* Closest hit for n rays (host arrays).  Returns WF_ERROR_STACK_OVERFLOW if
 * any ray overflowed its traversal stack (that ray reports a miss). */
WF_API wf_status wf_trace_closest(wf_renderer* r, const wf_ray* rays, uint32_t n, wf_hit* hits);
/* Same, on device arrays, ordered on the renderer's stream (no host sync).
 * A ray that overflows its traversal stack reports a miss, and the call
 * cannot return WF_ERROR_STACK_OVERFLOW without a sync: the overflow count of
 * this call is read by wf_renderer_read_stack_overflows. */
WF_API wf_status wf_trace_closest_device(wf_renderer* r, const wf_ray* d_rays, uint32_t n, wf_hit* d_hits);
/* Rays that overflowed the traversal stack in the last trace call or frame
 * (synchronises the renderer's stream).  No reference counterpart: the
 * reference's stacks are unchecked (accelstruct.hlsl:64,160). */
WF_API wf_status wf_renderer_read_stack_overflows(wf_renderer* r, uint32_t* out);
/* Occlusion: visible[i] = 1 if nothing is hit with t in [0, tmax). */
WF_API wf_status wf_trace_occluded(wf_renderer* r, const wf_ray* rays, uint32_t n, uint32_t* visible);
/* Per-ray traversal counters of the reference algorithm (4 x u32 per ray:
 * TLAS nodes fetched, BLAS nodes fetched, instances visited, triangles
 * tested), for the algorithmic-bytes roofline. */
WF_API wf_status wf_trace_closest_counters(wf_renderer* r, const wf_ray* rays, uint32_t n, wf_hit* hits, uint32_t* counters);

/* ---- host builders (bvh_builder_t / tlas_builder_t build + extract) ---- */
/* SAH binned BLAS over n triangles given as 9 floats each (p0,p1,p2).
 * Writes the reference's GPU blob (bvh_header_t ‖ nodes ‖ triangles ‖
 * indices, renderer.cpp:114-144) into out_blob when capacity suffices;
 * *out_size always receives the blob size, *out_node_count node_at. */
WF_API wf_status wf_bvh_build(const float* tri_positions, uint32_t n, const wf_build_options* options,
                              void* out_blob, uint64_t capacity, uint64_t* out_size, uint32_t* out_node_count);
/* The same BLAS built on the GPU (bvh_builder_t::build, bvh_builder.cpp:6-168,
 * run as a device build): d_tri_positions holds n triangles of 9 floats in
 * device memory; d_nodes receives the node array of the blob above
 * (node_capacity >= 2n nodes) and d_indices its index array (n entries), the
 * same bytes wf_bvh_build writes.  interval_count must be 2..64.  stream is a
 * hipStream_t (0 = null stream); the call returns after the build finished
 * and *out_node_count holds node_at. */
WF_API wf_status wf_bvh_build_device(const float* d_tri_positions, uint32_t n, const wf_build_options* options,
                                     wf_bvh_node* d_nodes, uint64_t node_capacity, uint32_t* d_indices,
                                     uint32_t* out_node_count, void* stream);
/* The same build as wf_bvh_build_device without allocating: no reference
 * counterpart; for callers that cannot allocate device memory (graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Its scratch is a
 * fixed device array of 32 MiB per GPU, held until the process ends, enough
 * for n <= 65536 triangles at any interval count: larger meshes return
 * WF_ERROR_CAPACITY.  Calls on one GPU share that array and are serialised
 * within a process.  Same arguments and results as wf_bvh_build_device. */
WF_API wf_status wf_bvh_build_device_noalloc(const float* d_tri_positions, uint32_t n, const wf_build_options* options,
                                             wf_bvh_node* d_nodes, uint64_t node_capacity, uint32_t* d_indices,
                                             uint32_t* out_node_count, void* stream);
/* A reusable device builder (bvh_builder_t, bvh/bvh_builder.h:14-37): the same
 * build as wf_bvh_build_device with its scratch kept across builds (it grows to
 * the largest mesh and is freed by destroy).  last_build reports the tree
 * levels built a level at a time, the subtrees built one per wave and the
 * largest leaf of the last build. */
typedef struct wf_bvh_builder wf_bvh_builder;
WF_API wf_status wf_bvh_builder_create(wf_bvh_builder** out_builder);
WF_API wf_status wf_bvh_builder_destroy(wf_bvh_builder* b);
WF_API wf_status wf_bvh_builder_build_device(wf_bvh_builder* b, const float* d_tri_positions, uint32_t n,
                                             const wf_build_options* options, wf_bvh_node* d_nodes, uint64_t node_capacity,
                                             uint32_t* d_indices, uint32_t* out_node_count, void* stream);
WF_API wf_status wf_bvh_builder_last_build(const wf_bvh_builder* b, uint32_t* levels, uint32_t* subtrees, uint32_t* max_leaf);
WF_API const char* wf_bvh_builder_last_error(const wf_bvh_builder* b);
/* Agglomerative TLAS blob (tlas_header_t ‖ nodes ‖ instances). */
WF_API wf_status wf_tlas_build(const wf_bvh_instance* instances, uint32_t m,
                               void* out_blob, uint64_t capacity, uint64_t* out_size, uint32_t* out_node_count);
/* The agglomerative walk's match metric (no reference counterpart for
 * UNION).  REFERENCE is tlas_builder.cpp:86-110's "max of the mins, max of
 * the maxes" (wf_tlas_build); UNION is the half area of the two boxes' union
 * (Walter et al. 2008), which gives a tree that rays cross in far fewer node
 * visits.  Closest hits are the same with either tree except where two
 * instances are hit at exactly the same distance, where the visit order
 * picks the one reported. */
typedef enum wf_tlas_metric { WF_TLAS_METRIC_REFERENCE = 0, WF_TLAS_METRIC_UNION = 1 } wf_tlas_metric;
WF_API wf_status wf_tlas_build_metric(const wf_bvh_instance* instances, uint32_t m, uint32_t metric,
                                      void* out_blob, uint64_t capacity, uint64_t* out_size, uint32_t* out_node_count);
/* The same TLAS built on the GPU (tlas_builder_t::build, tlas_builder.cpp:6-110):
 * d_instances holds m records in device memory; d_nodes (capacity >= 2m
 * nodes) receives the node array of the blob above, the same bytes
 * wf_tlas_build writes.  Returns after the build; *out_node_count = node count. */
WF_API wf_status wf_tlas_build_device(const wf_bvh_instance* d_instances, uint32_t m, wf_tlas_node* d_nodes,
                                      uint64_t node_capacity, uint32_t* out_node_count, void* stream);
/* wf_tlas_build_device with a choice of match metric (wf_tlas_metric). */
WF_API wf_status wf_tlas_build_device_metric(const wf_bvh_instance* d_instances, uint32_t m, uint32_t metric, wf_tlas_node* d_nodes,
                                             uint64_t node_capacity, uint32_t* out_node_count, void* stream);
/* The software instance record submit_render_mesh derives (renderer.cpp:1318-1336). */
WF_API wf_status wf_make_bvh_instance(const float transform[16], const float blas_min[3], const float blas_max[3],
                                      uint32_t bvh_index, wf_bvh_instance* out);

/* Self-check of the device arithmetic the traversal relies on (no reference
 * counterpart).  The traversal forms 1/x (ray and triangle reciprocals) as
 * the hardware reciprocal refined by one fused Newton step where |x| is in
 * [2^-126, 2^126], and as the IEEE division elsewhere; this compares that
 * value with the division 1.0f / x, bit for bit, on all 2^32 float inputs on
 * the current GPU (NaN results compare equal).  *out_mismatches receives the
 * count and *out_first_bits the bit pattern of the smallest mismatching input
 * (0xFFFFFFFF when none). */
WF_API wf_status wf_selftest_reciprocal(uint64_t* out_mismatches, uint32_t* out_first_bits);
/* Self-check of the cost-ordered queue (traversal config cost_order; no
 * reference counterpart): for n device rays and the box [bmin, bmax], writes
 * the queue order the trace calls would use (d_order, n device words: ray
 * indices, predicted-longest bucket first, callers' order within a bucket)
 * and each ray's cost bucket 0..15 (d_bucket, n device words).
 * Synchronises the device. */
WF_API wf_status wf_selftest_cost_order(const wf_ray* d_rays, uint32_t n, const float bmin[3], const float bmax[3],
                                        uint32_t* d_order, uint32_t* d_bucket);

#ifdef __cplusplus
}
#endif
#endif
