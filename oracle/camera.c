/*
 * Oracle: camera matrices and primary-ray generation.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

/* glm vec3 helpers (extern/glm/detail/func_geometric.inl): dot = (x+y)+z,
 * normalize = v * (1/sqrt(dot(v,v))). */
static float gdot(const float a[3], const float b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
static void gcross(const float x[3], const float y[3], float o[3])
{
    o[0] = x[1] * y[2] - y[1] * x[2];
    o[1] = x[2] * y[0] - y[2] * x[0];
    o[2] = x[0] * y[1] - y[0] * x[1];
}
static void gnormalize(const float v[3], float o[3])
{
    float s = 1.0f / sqrtf(gdot(v, v));
    o[0] = v[0] * s; o[1] = v[1] * s; o[2] = v[2] * s;
}

/* camera::create (core/camera/camera.cpp:8-16) -> glm::lookAtLH
 * (extern/glm/ext/matrix_transform.inl:176-196) with up = +y
 * (core/common.h:6); renderer::begin_scene (renderer.cpp:599-630) ->
 * perspectiveFovLH_ZO (extern/glm/ext/matrix_clip_space.inl:392-409). */
void orc_camera_view(const float eye[3], const float target[3], float vfov_deg,
                     uint32_t width, uint32_t height, orc_view* out)
{
    memset(out, 0, sizeof(*out));
    const float up[3] = { 0.0f, 1.0f, 0.0f };
    float fd[3] = { target[0] - eye[0], target[1] - eye[1], target[2] - eye[2] };
    float f[3], s0[3], s[3], u[3];
    gnormalize(fd, f);
    gcross(up, f, s0);
    gnormalize(s0, s);
    gcross(f, s, u);
    float* V = out->world_to_view;
    for (int i = 0; i < 16; ++i) V[i] = (i % 5 == 0) ? 1.0f : 0.0f;
    V[0 * 4 + 0] = s[0]; V[1 * 4 + 0] = s[1]; V[2 * 4 + 0] = s[2];
    V[0 * 4 + 1] = u[0]; V[1 * 4 + 1] = u[1]; V[2 * 4 + 1] = u[2];
    V[0 * 4 + 2] = f[0]; V[1 * 4 + 2] = f[1]; V[2 * 4 + 2] = f[2];
    V[3 * 4 + 0] = -gdot(s, eye);
    V[3 * 4 + 1] = -gdot(u, eye);
    V[3 * 4 + 2] = -gdot(f, eye);

    const float near_plane = 0.01f, far_plane = 1000.0f;
    float rad = vfov_deg * 0.01745329251994329576923690768489f;   /* glm::radians */
    float h = cosf(0.5f * rad) / sinf(0.5f * rad);
    float w = h * (float)height / (float)width;
    float* P = out->view_to_clip;
    memset(P, 0, 64);
    P[0 * 4 + 0] = w;
    P[1 * 4 + 1] = h;
    P[2 * 4 + 2] = far_plane / (far_plane - near_plane);
    P[2 * 4 + 3] = 1.0f;
    P[3 * 4 + 2] = -(far_plane * near_plane) / (far_plane - near_plane);

    orc_mat4_inverse(V, out->view_to_world);
    orc_mat4_inverse(P, out->clip_to_view);
    out->render_dim[0] = (float)width;
    out->render_dim[1] = (float)height;
    out->near_plane = near_plane;
    out->far_plane = far_plane;
}

/* generate.hlsl:12-20 with the software make_primary_ray / make_ray
 * (common.hlsl:145-169): the origin is nudged by dir * 1e-8f; the ray record
 * is a RayDesc2 with TMin = 1e-8f and TMax = FLT_MAX. */
void orc_primary_rays(const orc_view* v, uint32_t width, uint32_t height, orc_ray* rays)
{
    for (uint32_t py = 0; py < height; ++py) {
        for (uint32_t px = 0; px < width; ++px) {
            float uvx = ((float)px + 0.5f) / v->render_dim[0];
            float uvy = ((float)py + 0.5f) / v->render_dim[1];
            uvy = 1.0f - uvy;
            float c4[4] = { 2.0f * uvx - 1.0f, 2.0f * uvy - 1.0f, 1.0f, 1.0f };
            float cv[4];
            orc_mat4_mul_vec4(v->clip_to_view, c4, cv);
            float dv[3];
            gnormalize(cv, dv);
            float d4[4] = { dv[0], dv[1], dv[2], 0.0f }, dw[4];
            orc_mat4_mul_vec4(v->view_to_world, d4, dw);
            float o4[4] = { 0.0f, 0.0f, 0.0f, 1.0f }, ow[4];
            orc_mat4_mul_vec4(v->view_to_world, o4, ow);
            orc_ray* r = &rays[(size_t)py * width + px];
            for (int k = 0; k < 3; ++k) {
                r->origin[k] = ow[k] + dw[k] * 1e-8f;
                r->dir[k] = dw[k];
            }
            r->tmin = 1e-8f;
            r->tmax = 3.402823466e+38f;
        }
    }
}
