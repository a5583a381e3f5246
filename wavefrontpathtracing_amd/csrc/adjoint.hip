// Differentiable rendering entry points (no reference counterpart):
// path-replay adjoint of the last frame, in-place texture / emission
// updates for optimisation loops, and device readback of the image.
#include "renderer_impl.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>

using namespace wf;

static wf_status adjoint_run(wf_renderer* r, const float* d_dl, int32_t tex_slot, float* d_grad_tex, size_t tex_texels,
                             float* d_grad_em, size_t n_inst)
{
    if (!r->forward_valid)
        return r->fail(WF_ERROR_INVALID_STATE, "render_adjoint needs a render() of the same scene and settings just before it");
    if (r->settings.render_view_mode != 0)
        return r->fail(WF_ERROR_INVALID_ARGUMENT, "debug view modes have no adjoint");
    FrameArgs A = r->last_args;
    const size_t n_paths = r->frame_capacity;
    hipError_t e = hipSuccess;
    auto need = [&](DevBuf& b, size_t bytes) { if (e == hipSuccess) e = b.ensure(std::max<size_t>(bytes, 64)); };
    need(r->d_lrem, 16 * n_paths);
    need(r->d_delta, 16 * n_paths);
    need(r->d_shadow_unit, 16 * n_paths);
    need(r->d_shadow_light, 4 * n_paths);
    if (e != hipSuccess) return r->hip_fail(e, "adjoint buffer allocation");
    A.lrem = (float4*)r->d_lrem.p;
    A.delta = (float4*)r->d_delta.p;
    A.dl_dimage = (const float4*)d_dl;
    A.shadow_unit = (float4*)r->d_shadow_unit.p;
    A.shadow_light = (uint32_t*)r->d_shadow_light.p;
    A.grad_tex = (float4*)d_grad_tex;
    A.grad_em = d_grad_em;
    A.grad_tex_slot = tex_slot;
    const uint32_t grid = (uint32_t)r->cu_count * 8u;
    wf_status s;
    if ((e = hipEventRecord(r->ev[0], r->stream)) != hipSuccess) return r->hip_fail(e, "event");
    // in chunks of 2^28 words: a fill's count is 32-bit
    for (size_t off = 0, words = 4 * tex_texels, chunk = size_t(1) << 28; off < words; off += chunk)
        if ((e = launch_fill_u32((uint32_t*)d_grad_tex + off, 0u, (uint32_t)std::min(chunk, words - off), r->stream)) != hipSuccess)
            return r->hip_fail(e, "gradient clear");
    if ((e = launch_fill_u32((uint32_t*)d_grad_em, 0u, (uint32_t)n_inst, r->stream)) != hipSuccess) return r->hip_fail(e, "gradient clear");
    if ((s = r->reset_overflows()) != WF_OK) return s;
    if ((e = launch_adjoint_clear(A, grid, r->stream)) != hipSuccess) return r->hip_fail(e, "adjoint clear launch");
    if ((e = launch_generate(A, grid, r->stream)) != hipSuccess) return r->hip_fail(e, "generate launch");
    const bool batched = r->slots.size() > 1;
    // as in the forward frame, bounce d-1's shadow rays may ride in bounce d's
    // extend launch: adjoint_connect(d-1) then runs before adjoint_shade(d),
    // the order in which the separate launches update the remaining estimate
    const bool fuse = A.nee && !r->settings.separate_connect && (r->ext_cfg.skip_popped || r->ext_cfg.layout == 0) &&
                      r->ext_grid_tlds == 0;
    for (uint32_t d = 0; d <= A.max_bounces; ++d) {
        const bool with_shadow = fuse && d > 0;
        wf_renderer::OcclusionQueue q{A.shadow_rays, A.counts + kShadowCountBase + (d > 0 ? d - 1 : 0), A.shadow_visible,
                         batched ? A.shadow_root : nullptr};
        s = r->run_extend(A.rays[d & 1u], A.counts + d, A.n_paths, A.hits, nullptr, nullptr, false,
                          batched ? A.root_of[d & 1u] : nullptr, with_shadow ? &q : nullptr);
        if (s != WF_OK) return s;
        if (with_shadow && (e = launch_adjoint_connect(A, d - 1, grid, r->stream)) != hipSuccess)
            return r->hip_fail(e, "adjoint connect launch");
        if ((e = launch_adjoint_shade(A, d, grid, r->stream)) != hipSuccess) return r->hip_fail(e, "adjoint shade launch");
        if (A.nee && !fuse && d < A.max_bounces) {
            s = r->run_extend(A.shadow_rays, A.counts + kShadowCountBase + d, A.n_paths, nullptr, A.shadow_visible, nullptr,
                              true, batched ? A.shadow_root : nullptr);
            if (s != WF_OK) return s;
            if ((e = launch_adjoint_connect(A, d, grid, r->stream)) != hipSuccess) return r->hip_fail(e, "adjoint connect launch");
        }
    }
    if ((e = hipEventRecord(r->ev[1], r->stream)) != hipSuccess) return r->hip_fail(e, "event");
    if ((e = hipStreamSynchronize(r->stream)) != hipSuccess) return r->hip_fail(e, "adjoint");
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, r->ev[0], r->ev[1]);
    r->stats.ms_total = ms;
    uint32_t ov = 0;
    if ((s = r->read_overflows(&ov)) != WF_OK) return s;
    if (ov) return r->fail(WF_ERROR_STACK_OVERFLOW, "%u rays overflowed the traversal stack in the adjoint", ov);
    return WF_OK;
}

static wf_status adjoint_shape(wf_renderer* r, wf_texture_handle tex, int32_t* slot, size_t* texels, size_t* n_inst, size_t* n_pix)
{
    *slot = -1;
    *texels = 0;
    if (tex != WF_INVALID_HANDLE) {
        Texture* t = r->find_texture(tex);
        if (!t) return r->fail(WF_ERROR_NOT_FOUND, "texture handle is not valid");
        *slot = (int32_t)(tex & 0xFFFFFFFFu);
        *texels = (size_t)t->width * t->height;
        // the replay differentiates the albedo use of the texture only: one
        // that also supplies emission or the environment would miss terms
        for (const auto& sl : r->slots)
            for (const auto& rec : sl.instances)
                if (rec.material.emissive_texture == tex)
                    return r->fail(WF_ERROR_INVALID_ARGUMENT, "the differentiated texture is also an emissive texture");
        if (r->env_texture == tex)
            return r->fail(WF_ERROR_INVALID_ARGUMENT, "the differentiated texture is also the environment texture");
    }
    *n_inst = 0;
    for (const auto& sl : r->slots) *n_inst += sl.instances.size();
    *n_pix = r->image_pixels;
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_render_adjoint_device(wf_renderer* r, const float* d_dl_dimage, wf_texture_handle texture,
                                                             float* d_grad_texels, float* d_grad_emission)
{
    if (!r || !d_dl_dimage || !d_grad_emission) return WF_ERROR_INVALID_ARGUMENT;
    int32_t slot;
    size_t texels, n_inst, n_pix;
    wf_status s = adjoint_shape(r, texture, &slot, &texels, &n_inst, &n_pix);
    if (s != WF_OK) return s;
    if (slot >= 0 && !d_grad_texels) return WF_ERROR_INVALID_ARGUMENT;
    return adjoint_run(r, d_dl_dimage, slot, d_grad_texels, texels, d_grad_emission, n_inst);
}

extern "C" WF_API wf_status wf_renderer_render_adjoint(wf_renderer* r, const float* dl_dimage, wf_texture_handle texture,
                                                      float* out_grad_texels, float* out_grad_emission)
{
    if (!r || !dl_dimage || !out_grad_emission) return WF_ERROR_INVALID_ARGUMENT;
    int32_t slot;
    size_t texels, n_inst, n_pix;
    wf_status s = adjoint_shape(r, texture, &slot, &texels, &n_inst, &n_pix);
    if (s != WF_OK) return s;
    if (slot >= 0 && !out_grad_texels) return WF_ERROR_INVALID_ARGUMENT;
    hipError_t e = r->d_adj_dl.ensure(std::max<size_t>(16 * n_pix, 64));
    if (e == hipSuccess) e = r->d_adj_tex.ensure(std::max<size_t>(16 * texels, 64));
    if (e == hipSuccess) e = r->d_adj_em.ensure(std::max<size_t>(4 * n_inst, 64));
    if (e == hipSuccess) e = hipMemcpyAsync(r->d_adj_dl.p, dl_dimage, 16 * n_pix, hipMemcpyHostToDevice, r->stream);
    if (e != hipSuccess) return r->hip_fail(e, "adjoint upload");
    s = adjoint_run(r, (const float*)r->d_adj_dl.p, slot, (float*)r->d_adj_tex.p, texels, (float*)r->d_adj_em.p, n_inst);
    if (s != WF_OK) return s;
    if (slot >= 0) e = hipMemcpyAsync(out_grad_texels, r->d_adj_tex.p, 16 * texels, hipMemcpyDeviceToHost, r->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out_grad_emission, r->d_adj_em.p, 4 * n_inst, hipMemcpyDeviceToHost, r->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) return r->hip_fail(e, "adjoint readback");
    return WF_OK;
}

static wf_status update_texture(wf_renderer* r, wf_texture_handle tex, const float* rgba, hipMemcpyKind kind)
{
    if (!r || !rgba) return WF_ERROR_INVALID_ARGUMENT;
    Texture* t = r->find_texture(tex);
    if (!t) return r->fail(WF_ERROR_NOT_FOUND, "texture handle is not valid");
    hipError_t e = hipMemcpyAsync(t->texels.p, rgba, 16ull * t->width * t->height, kind, r->stream);
    if (e == hipSuccess && kind == hipMemcpyHostToDevice) e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) return r->hip_fail(e, "texture update");
    r->forward_valid = false;
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_update_texture(wf_renderer* r, wf_texture_handle tex, const float* rgba)
{
    return update_texture(r, tex, rgba, hipMemcpyHostToDevice);
}

extern "C" WF_API wf_status wf_renderer_update_texture_device(wf_renderer* r, wf_texture_handle tex, const float* d_rgba)
{
    return update_texture(r, tex, d_rgba, hipMemcpyDeviceToDevice);
}

extern "C" WF_API wf_status wf_renderer_set_emission_strength(wf_renderer* r, uint32_t instance, float strength)
{
    if (!r || !(strength >= 0.0f) || !std::isfinite(strength)) return WF_ERROR_INVALID_ARGUMENT;
    uint32_t base = 0;
    for (auto& sl : r->slots) {
        if (instance < base + sl.instances.size()) {
            InstanceRec& rec = sl.instances[instance - base];
            if (rec.material.emissive_strength == strength) return WF_OK;
            // an instance is an emitter (ends paths, sits in the emitter table)
            // by its submitted material; a strength update keeps it one
            const float* f = rec.material.emissive_factor;
            auto emits = [f](float s) { return s * f[0] > 0.0f || s * f[1] > 0.0f || s * f[2] > 0.0f; };
            if (!emits(rec.material.emissive_strength) || !emits(strength))
                return r->fail(WF_ERROR_INVALID_ARGUMENT, "instance %u: emission strength can only change to another positive "
                               "value on an instance submitted as emissive", instance);
            rec.material.emissive_strength = strength;
            if (r->tlas_valid) {
                char* dst = (char*)r->d_shade_inst.p + sizeof(GpuShadeInstance) * instance + offsetof(GpuShadeInstance, emissive_strength);
                hipError_t e = hipMemcpyAsync(dst, &strength, 4, hipMemcpyHostToDevice, r->stream);
                if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
                if (e != hipSuccess) return r->hip_fail(e, "emission update");
            }
            r->forward_valid = false;
            return WF_OK;
        }
        base += (uint32_t)sl.instances.size();
    }
    return r->fail(WF_ERROR_INVALID_ARGUMENT, "instance %u out of range", instance);
}

extern "C" WF_API wf_status wf_renderer_read_energy_device(wf_renderer* r, float* d_out)
{
    if (!r || !d_out) return WF_ERROR_INVALID_ARGUMENT;
    if (!r->image_valid) return r->fail(WF_ERROR_INVALID_STATE, "no frame rendered");
    hipError_t e = hipMemcpyAsync(d_out, r->d_image.p, 16ull * r->image_pixels, hipMemcpyDeviceToDevice, r->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) return r->hip_fail(e, "image copy");
    return WF_OK;
}
