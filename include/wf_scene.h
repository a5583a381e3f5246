/*
 * wf_scene.h — procedural inputs for the benchmark and parity workloads.
 *
 * No reference counterpart: the reference loads glTF/FBX assets
 * (core/assets/asset_loader.cpp:937) and hands their de-indexed triangles to
 * renderer::create_render_mesh.  These generators produce the same kind of
 * input (wf_vertex arrays + indices, local_to_world transforms) from a seeded
 * PCG32 (O'Neill 2014, "PCG: A Family of Simple Fast Space-Efficient
 * Statistically Good Algorithms for Random Number Generation"), so every
 * workload is reproducible bit for bit on any x86-64 host.
 *
 * Two-call protocol: pass NULL outputs to receive the counts.
 */
#ifndef WF_SCENE_H
#define WF_SCENE_H

#include "wf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wfs_pcg32 { uint64_t state; uint64_t inc; } wfs_pcg32;
WF_API void     wfs_pcg32_seed(wfs_pcg32* rng, uint64_t initstate, uint64_t initseq);
WF_API uint32_t wfs_pcg32_next(wfs_pcg32* rng);
/* uniform in [0,1): (next >> 8) * 2^-24 */
WF_API float    wfs_pcg32_float(wfs_pcg32* rng);

/* n triangles; centroid uniform in [-extent, extent]^3; each vertex =
 * centroid + per-component random sign * U(off_min, off_max).  Flat vertex
 * normals along (v1-v0) x (v2-v0); uvs (0,0) (1,0) (0,1).  3n vertices and
 * 3n indices (identity). */
WF_API wf_status wfs_random_soup(uint64_t seed, uint32_t n, float extent, float off_min, float off_max,
                                 wf_vertex* vertices, uint32_t* indices);
/* nx*nz quads over [-1,1]^2 in xz, y = amplitude * value noise (lattice 8x8,
 * smoothstep-bilinear); front faces point +y; flat normals; uv = grid coords
 * in [0,1].  (nx+1)(nz+1) vertices, 6 nx nz indices. */
WF_API wf_status wfs_heightfield(uint64_t seed, uint32_t nx, uint32_t nz, float amplitude,
                                 wf_vertex* vertices, uint32_t* indices);
/* unit icosphere, subdiv levels of 4-way midpoint subdivision: 20*4^subdiv
 * triangles, outward fronts, smooth normals, spherical uvs; unindexed
 * (3 vertices per triangle). */
WF_API wf_status wfs_icosphere(uint32_t subdiv, wf_vertex* vertices, uint32_t* indices,
                               uint32_t* out_vertex_count, uint32_t* out_index_count);
/* closed room [-1,1]^3 seen from inside: 6 walls, each split into 2 quads
 * (12 quads, 24 triangles), fronts facing inwards.  48 vertices, 72 indices. */
WF_API wf_status wfs_box_room(wf_vertex* vertices, uint32_t* indices);
/* unit quad [-0.5,0.5]^2 in the y=0 plane, front facing -y (an area light
 * pointing down).  4 vertices, 6 indices. */
WF_API wf_status wfs_quad(wf_vertex* vertices, uint32_t* indices);
/* local_to_world = T * R * S: R uniform random rotation (Shoemake 1992),
 * S uniform scale U(scale_min, scale_max), T uniform in [-t_extent, t_extent]^3.
 * Column-major. */
WF_API void wfs_random_transform(wfs_pcg32* rng, float scale_min, float scale_max, float t_extent, float out[16]);
/* n rays: origin uniform in [bmin,bmax], direction uniform on the sphere,
 * tmin 1e-8, tmax FLT_MAX. */
WF_API wf_status wfs_sphere_rays(uint64_t seed, const float bmin[3], const float bmax[3], uint32_t n, wf_ray* out);
/* the camera rays of generate (make_primary_ray, common.hlsl:145-169) for a
 * w x h image of view v, row-major (ray y*w + x), tmin 1e-8, tmax FLT_MAX */
WF_API wf_status wfs_primary_rays(const wf_view* v, uint32_t width, uint32_t height, wf_ray* out);
/* w x h RGB float texture of periodic value noise (lattice period `cells`)
 * mapped into [lo, hi]. */
WF_API wf_status wfs_value_noise_texture(uint64_t seed, uint32_t w, uint32_t h, uint32_t cells, float lo, float hi, float* out_rgb);

#ifdef __cplusplus
}
#endif
#endif
