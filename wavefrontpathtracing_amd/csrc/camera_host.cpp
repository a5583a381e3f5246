// Camera and view constants (host only): camera::create
// (core/camera/camera.cpp:8-16) and the view_t that renderer::begin_scene
// derives (renderer.cpp:599-630), with glm's float arithmetic.
#include "wf_host.h"
#include "../../include/wf_scene.h"

#include <cfloat>
#include <cmath>
#include <cstring>

namespace wf {

void make_view(const wf_camera& cam, uint32_t width, uint32_t height, wf_view& v)
{
    std::memset(&v, 0, sizeof(v));
    const float near_plane = 0.01f, far_plane = 1000.0f;
    // glm::perspectiveFovLH_ZO (extern/glm/ext/matrix_clip_space.inl:392-409)
    const float rad = cam.vfov_deg * 0.01745329251994329576923690768489f;   // glm::radians
    const float h = std::cos(0.5f * rad) / std::sin(0.5f * rad);
    const float w = h * (float)height / (float)width;
    float* P = v.view_to_clip;
    P[0] = w;
    P[5] = h;
    P[10] = far_plane / (far_plane - near_plane);
    P[11] = 1.0f;
    P[14] = -(far_plane * near_plane) / (far_plane - near_plane);
    std::memcpy(v.world_to_view, cam.view_matrix, 64);
    mat4_inverse(cam.view_matrix, v.view_to_world);
    mat4_inverse(v.view_to_clip, v.clip_to_view);
    v.render_dim[0] = (float)width;
    v.render_dim[1] = (float)height;
    v.near_plane = near_plane;
    v.far_plane = far_plane;
}

} // namespace wf

extern "C" WF_API wf_status wf_camera_create(wf_camera* out, const float eye[3], const float target[3], float vfov_deg)
{
    if (!out || !eye || !target || !(vfov_deg > 0.0f && vfov_deg < 180.0f)) return WF_ERROR_INVALID_ARGUMENT;
    std::memset(out, 0, sizeof(*out));
    out->vfov_deg = vfov_deg;
    // glm::lookAtLH with up = +y (extern/glm/ext/matrix_transform.inl:176-196)
    const float up[3] = {0.0f, 1.0f, 0.0f};
    const float fd[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
    float f[3], s0[3], s[3], u[3];
    wf::normalize3(fd, f);
    wf::cross3(up, f, s0);
    wf::normalize3(s0, s);
    wf::cross3(f, s, u);
    float* V = out->view_matrix;
    for (int i = 0; i < 16; ++i) V[i] = (i % 5 == 0) ? 1.0f : 0.0f;
    for (int k = 0; k < 3; ++k) { V[k * 4 + 0] = s[k]; V[k * 4 + 1] = u[k]; V[k * 4 + 2] = f[k]; }
    V[12] = -wf::dot3(s, eye);
    V[13] = -wf::dot3(u, eye);
    V[14] = -wf::dot3(f, eye);
    wf::mat4_inverse(V, out->inv_view_matrix);
    return WF_OK;
}


extern "C" WF_API wf_status wf_make_view(const wf_camera* camera, uint32_t width, uint32_t height, wf_view* out)
{
    if (!camera || !out || width == 0 || height == 0 || !(camera->vfov_deg > 0.0f && camera->vfov_deg < 180.0f))
        return WF_ERROR_INVALID_ARGUMENT;
    wf::make_view(*camera, width, height, *out);
    return WF_OK;
}

// The camera rays the generate stage makes (make_primary_ray + make_ray,
// common.hlsl:145-169; generate.hlsl:12-20), on the host: pixel (x, y) of a
// w x h image is ray y * w + x.
extern "C" WF_API wf_status wfs_primary_rays(const wf_view* v, uint32_t width, uint32_t height, wf_ray* out)
{
    if (!v || !out || width == 0 || height == 0) return WF_ERROR_INVALID_ARGUMENT;
    const float zero4[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    float ow[4];
    wf::mat4_mul_vec4(v->view_to_world, zero4, ow);
    for (uint32_t py = 0; py < height; ++py) {
        for (uint32_t px = 0; px < width; ++px) {
            const float uvx = ((float)px + 0.5f) / v->render_dim[0];
            float uvy = ((float)py + 0.5f) / v->render_dim[1];
            uvy = 1.0f - uvy;
            const float c4[4] = {2.0f * uvx - 1.0f, 2.0f * uvy - 1.0f, 1.0f, 1.0f};
            float cv[4], dv[3], d4[4], dw[4];
            wf::mat4_mul_vec4(v->clip_to_view, c4, cv);
            wf::normalize3(cv, dv);
            d4[0] = dv[0]; d4[1] = dv[1]; d4[2] = dv[2]; d4[3] = 0.0f;
            wf::mat4_mul_vec4(v->view_to_world, d4, dw);
            wf_ray& r = out[(size_t)py * width + px];
            for (int k = 0; k < 3; ++k) {
                r.origin[k] = ow[k] + dw[k] * 1e-8f;
                r.direction[k] = dw[k];
            }
            r.tmin = 1e-8f;
            r.tmax = FLT_MAX;
        }
    }
    return WF_OK;
}
