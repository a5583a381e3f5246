// Frame / scene lifecycle (renderer.cpp:543-630, 1051-1055) and the
// material / emitter tables the shade and connect stages read.
#include "renderer_impl.h"

#include <cfloat>
#include <cmath>
#include <cstring>

using namespace wf;

extern "C" WF_API wf_status wf_renderer_begin_frame(wf_renderer* r)
{
    if (!r) return WF_ERROR_INVALID_ARGUMENT;
    if (r->frame_open) return r->fail(WF_ERROR_INVALID_STATE, "begin_frame called twice");
    r->frame_open = true;
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_end_frame(wf_renderer* r)
{
    if (!r) return WF_ERROR_INVALID_ARGUMENT;
    if (!r->frame_open) return r->fail(WF_ERROR_INVALID_STATE, "end_frame without begin_frame");
    r->frame_open = false;
    r->frame_index++;
    return WF_OK;
}

// begin_scene (renderer.cpp:599-630)
extern "C" WF_API wf_status wf_renderer_begin_scene(wf_renderer* r, const wf_camera* cam, wf_texture_handle env)
{
    if (!r || !cam) return WF_ERROR_INVALID_ARGUMENT;
    if (r->scene_open) return r->fail(WF_ERROR_INVALID_STATE, "begin_scene called twice");
    if (!(cam->vfov_deg > 0.0f && cam->vfov_deg < 180.0f)) return r->fail(WF_ERROR_INVALID_ARGUMENT, "vfov_deg out of range");
    if (std::memcmp(cam->view_matrix, r->camera.view_matrix, 64) != 0) r->accum_count = 0;   // reset_color_accumulator
    r->camera = *cam;
    r->env_texture = env;
    make_view(*cam, r->width, r->height, r->view);
    r->view_valid = true;
    r->scene_open = true;
    r->instances.clear();
    r->tlas_valid = false;
    return WF_OK;
}

// end_scene (renderer.cpp:1051-1055): instances live for one scene
extern "C" WF_API wf_status wf_renderer_end_scene(wf_renderer* r)
{
    if (!r) return WF_ERROR_INVALID_ARGUMENT;
    if (!r->scene_open) return r->fail(WF_ERROR_INVALID_STATE, "end_scene without begin_scene");
    r->scene_open = false;
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_get_view(const wf_renderer* r, wf_view* out)
{
    if (!r || !out) return WF_ERROR_INVALID_ARGUMENT;
    if (!r->view_valid) return WF_ERROR_INVALID_STATE;
    *out = r->view;
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_get_frame_stats(const wf_renderer* r, wf_frame_stats* out)
{
    if (!r || !out) return WF_ERROR_INVALID_ARGUMENT;
    *out = r->stats;
    return WF_OK;
}

void wf_renderer::release_frame_buffers()
{
    for (auto& b : d_rays) b.release();
    for (auto& b : d_paths) b.release();
    d_hits.release(); d_shadow.release(); d_counts.release(); d_energy.release(); d_accum.release(); d_final.release();
    frame_capacity = 0;
}

// Texture table and the emitter list for next-event estimation: every
// triangle of an instance whose emissive factor is non-zero, with its world
// area (local triangle transformed by local_to_world).
wf_status wf_renderer::upload_materials_and_emitters()
{
    std::vector<GpuTexture> tt(std::max<size_t>(textures.size(), 1));
    for (size_t i = 0; i < textures.size(); ++i) {
        tt[i].texels = (const float*)textures[i].texels.p;
        tt[i].width = textures[i].width;
        tt[i].height = textures[i].height;
    }
    hipError_t e = d_textures.ensure(sizeof(GpuTexture) * tt.size());
    if (e == hipSuccess) e = hipMemcpy(d_textures.p, tt.data(), sizeof(GpuTexture) * tt.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "texture table upload");

    // emitters: (instance, primitive, cdf) -- uniform by world area
    std::vector<uint32_t> ids;
    std::vector<double> area;
    for (size_t i = 0; i < instances.size(); ++i) {
        const InstanceRec& rec = instances[i];
        const float em = rec.material.emissive_strength;
        const float* ef = rec.material.emissive_factor;
        if (!(em * ef[0] > 0.0f || em * ef[1] > 0.0f || em * ef[2] > 0.0f)) continue;
        const Mesh& m = meshes[rec.mesh];
        for (uint32_t t = 0; t < m.tri_count; ++t) {
            float w[3][4];
            for (int v = 0; v < 3; ++v) {
                const float p[4] = {m.pos9[9ull * t + 3 * v], m.pos9[9ull * t + 3 * v + 1], m.pos9[9ull * t + 3 * v + 2], 1.0f};
                mat4_mul_vec4(rec.transform, p, w[v]);
            }
            const double e1[3] = {(double)w[1][0] - w[0][0], (double)w[1][1] - w[0][1], (double)w[1][2] - w[0][2]};
            const double e2[3] = {(double)w[2][0] - w[0][0], (double)w[2][1] - w[0][1], (double)w[2][2] - w[0][2]};
            const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
            const double a = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
            if (!(a > 0.0)) continue;
            ids.push_back((uint32_t)i);
            ids.push_back(t);
            area.push_back(a);
        }
    }
    n_emitters = (uint32_t)area.size();
    double total = 0.0;
    for (double a : area) total += a;
    emitter_area = (float)total;
    // record: (instance, primitive, cdf as float bits, pad)
    std::vector<uint32_t> rec(4 * std::max<size_t>(area.size(), 1), 0u);
    double acc = 0.0;
    for (size_t k = 0; k < area.size(); ++k) {
        acc += area[k];
        float c = (float)(acc / total);
        if (k + 1 == area.size()) c = 1.0f;
        rec[4 * k] = ids[2 * k];
        rec[4 * k + 1] = ids[2 * k + 1];
        std::memcpy(&rec[4 * k + 2], &c, 4);
    }
    e = d_emitters.ensure(4 * rec.size());
    if (e == hipSuccess) e = hipMemcpy(d_emitters.p, rec.data(), 4 * rec.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "emitter table upload");
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_render(wf_renderer* r)
{
    if (!r) return WF_ERROR_INVALID_ARGUMENT;
    return r->fail(WF_ERROR_INVALID_STATE, "render: wavefront loop not built yet");
}

extern "C" WF_API wf_status wf_renderer_read_energy(wf_renderer* r, float* out)
{
    if (!r || !out) return WF_ERROR_INVALID_ARGUMENT;
    return r->fail(WF_ERROR_INVALID_STATE, "no frame rendered");
}

extern "C" WF_API wf_status wf_renderer_read_accumulator(wf_renderer* r, float* out)
{
    if (!r || !out) return WF_ERROR_INVALID_ARGUMENT;
    return r->fail(WF_ERROR_INVALID_STATE, "no frame rendered");
}

extern "C" WF_API wf_status wf_renderer_read_final_color(wf_renderer* r, uint8_t* out)
{
    if (!r || !out) return WF_ERROR_INVALID_ARGUMENT;
    return r->fail(WF_ERROR_INVALID_STATE, "no frame rendered");
}
