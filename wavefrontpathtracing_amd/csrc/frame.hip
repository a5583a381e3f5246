// Frame / scene lifecycle (renderer.cpp:543-630, 1051-1055) and the
// material / emitter tables the shade and connect stages read.
#include "renderer_impl.h"
#include "stages.h"

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
    d_hits.release(); d_thr.release(); d_path_energy.release(); d_counts.release();
    d_shadow_rays.release(); d_shadow_path.release(); d_shadow_contrib.release(); d_shadow_visible.release();
    d_image.release(); d_accum.release(); d_final.release();
    frame_capacity = 0;
    image_valid = false;
}

// Queues sized for every path of the frame (the reference sizes them to
// W*H, renderer.cpp:432-469); the accumulator and output images per pixel.
wf_status wf_renderer::ensure_frame_buffers(size_t n_paths, size_t n_pixels)
{
    hipError_t e = hipSuccess;
    auto need = [&](DevBuf& b, size_t bytes) { if (e == hipSuccess) e = b.ensure(std::max<size_t>(bytes, 64)); };
    for (auto& b : d_rays) need(b, sizeof(wf_ray) * n_paths);
    for (auto& b : d_paths) need(b, 4 * n_paths);
    need(d_hits, sizeof(wf_hit) * n_paths);
    need(d_thr, 16 * n_paths);
    need(d_path_energy, 16 * n_paths);
    need(d_counts, 4 * kCountSlots);
    need(d_shadow_rays, sizeof(wf_ray) * n_paths);
    need(d_shadow_path, 4 * n_paths);
    need(d_shadow_contrib, 16 * n_paths);
    need(d_shadow_visible, 4 * n_paths);
    const bool fresh_accum = d_accum.bytes < 16 * n_pixels;
    need(d_image, 16 * n_pixels);
    need(d_accum, 16 * n_pixels);
    need(d_final, 4 * n_pixels);
    if (e != hipSuccess) return hip_fail(e, "frame buffer allocation");
    if (fresh_accum) accum_count = 0;
    frame_capacity = n_paths;
    return WF_OK;
}

// renderer::render (renderer.cpp:632-1049), wavefront branch: clear, then
// for every depth d <= max_bounces: (generate at d = 0), extend, shade,
// connect; then resolve + post-process.  All counts stay on the device.
wf_status wf_renderer::render_frame()
{
    if (!view_valid) return fail(WF_ERROR_INVALID_STATE, "render called before begin_scene");
    if (env_texture != WF_INVALID_HANDLE && find_texture(env_texture))
        return fail(WF_ERROR_INVALID_ARGUMENT, "environment textures are not supported (only the constant default)");
    wf_status s = upload_scene();
    if (s != WF_OK) return s;
    const uint32_t npix = width * height;
    const uint64_t npath64 = (uint64_t)npix * settings.spp;
    if (npath64 > (1ull << 30)) return fail(WF_ERROR_CAPACITY, "width*height*spp exceeds 2^30 paths");
    const uint32_t npath = (uint32_t)npath64;
    if ((s = ensure_frame_buffers(npath, npix)) != WF_OK) return s;
    if (accum_count == 0) accum_count = 1;   // reset_color_accumulator (renderer.cpp:47-51)

    // renderer.cpp:739: a new frame seed per frame unless the caller fixes it
    uint32_t seed = settings.frame_seed;
    if (seed == 0) {
        xorshift_seed ^= xorshift_seed << 13;
        xorshift_seed ^= xorshift_seed >> 17;
        xorshift_seed ^= xorshift_seed << 5;
        seed = xorshift_seed;
    }

    FrameArgs A{};
    A.width = width; A.height = height; A.spp = settings.spp; A.n_paths = npath;
    A.view = view;
    A.frame_seed = seed;
    A.max_bounces = settings.max_bounces;
    A.rr_start = settings.rr_start_bounce;
    A.nee = settings.next_event_estimation;
    A.rr_survival = settings.rr_survival;
    for (int k = 0; k < 3; ++k) A.env[k] = settings.hdr_env_strength;   // default 1x1 white environment
    A.n_emitters = n_emitters;
    A.emitter_area = emitter_area;
    A.accumulate = settings.accumulate;
    A.sample_count = accum_count;
    A.thr = (float4*)d_thr.p;
    A.energy = (float4*)d_path_energy.p;
    for (int k = 0; k < 2; ++k) { A.rays[k] = (wf_ray*)d_rays[k].p; A.path_of[k] = (uint32_t*)d_paths[k].p; }
    A.hits = (wf_hit*)d_hits.p;
    A.counts = (uint32_t*)d_counts.p;
    A.shadow_rays = (wf_ray*)d_shadow_rays.p;
    A.shadow_path = (uint32_t*)d_shadow_path.p;
    A.shadow_contrib = (float4*)d_shadow_contrib.p;
    A.shadow_visible = (uint32_t*)d_shadow_visible.p;
    A.shade_inst = (const GpuShadeInstance*)d_shade_inst.p;
    A.textures = (const GpuTexture*)d_textures.p;
    A.emitters = (const uint4*)d_emitters.p;
    A.image = (float4*)d_image.p;
    A.accum = (float4*)d_accum.p;
    A.final_rgba8 = (uint32_t*)d_final.p;

    const uint32_t grid = (uint32_t)cu_count * 8u;
    hipError_t e;
    int ei = 0;
    auto mark = [&]() { return hipEventRecord(ev[ei++], stream); };
    if ((s = reset_overflows()) != WF_OK) return s;
    if ((e = mark()) != hipSuccess) return hip_fail(e, "event");
    if ((e = launch_clear(A, grid, stream)) != hipSuccess) return hip_fail(e, "clear launch");
    if ((e = mark()) != hipSuccess) return hip_fail(e, "event");
    if ((e = launch_generate(A, grid, stream)) != hipSuccess) return hip_fail(e, "generate launch");
    if ((e = mark()) != hipSuccess) return hip_fail(e, "event");
    // events: 0 start, 1 after clear, 2 after generate, then per depth
    // (after extend, after shade, after connect), then after post
    for (uint32_t d = 0; d <= settings.max_bounces; ++d) {
        s = run_extend(A.rays[d & 1u], A.counts + d, npath, A.hits, nullptr, nullptr, false);
        if (s != WF_OK) return s;
        if ((e = mark()) != hipSuccess) return hip_fail(e, "event");
        if ((e = launch_shade(A, d, grid, stream)) != hipSuccess) return hip_fail(e, "shade launch");
        if ((e = mark()) != hipSuccess) return hip_fail(e, "event");
        if (settings.next_event_estimation && d < settings.max_bounces) {
            s = run_extend(A.shadow_rays, A.counts + kShadowCountBase + d, npath, nullptr, A.shadow_visible, nullptr, true);
            if (s != WF_OK) return s;
            if ((e = launch_connect_accumulate(A, d, grid, stream)) != hipSuccess) return hip_fail(e, "connect launch");
        }
        if ((e = mark()) != hipSuccess) return hip_fail(e, "event");
    }
    if ((e = launch_resolve_post(A, grid, stream)) != hipSuccess) return hip_fail(e, "post-process launch");
    if ((e = mark()) != hipSuccess) return hip_fail(e, "event");

    uint32_t counts[kCountSlots];
    e = hipMemcpyAsync(counts, d_counts.p, sizeof(counts), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "frame");
    wf_frame_stats st{};
    auto ms = [&](int a, int b) { float t = 0.0f; (void)hipEventElapsedTime(&t, ev[a], ev[b]); return t; };
    st.ms_clear = ms(0, 1);
    st.ms_generate = ms(1, 2);
    for (uint32_t d = 0; d <= settings.max_bounces; ++d) {
        const int b = 2 + 3 * (int)d;
        st.ms_extend += ms(b, b + 1);
        st.ms_shade += ms(b + 1, b + 2);
        st.ms_connect += ms(b + 2, b + 3);
        st.rays_extended += counts[d];
        if (d < kMaxBounces) st.shadow_rays += counts[kShadowCountBase + d];
    }
    st.ms_post = ms(ei - 2, ei - 1);
    st.ms_total = ms(0, ei - 1);
    uint32_t ov = 0;
    if ((s = read_overflows(&ov)) != WF_OK) return s;
    st.stack_overflows = ov;
    stats = st;
    image_valid = true;
    accum_count++;
    if (ov) return fail(WF_ERROR_STACK_OVERFLOW, "%u rays overflowed the traversal stack this frame", ov);
    return WF_OK;
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
    return r->render_frame();
}

static wf_status read_image(wf_renderer* r, const DevBuf& b, void* out, size_t bytes)
{
    if (!r->image_valid) return r->fail(WF_ERROR_INVALID_STATE, "no frame rendered");
    hipError_t e = hipMemcpyAsync(out, b.p, bytes, hipMemcpyDeviceToHost, r->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) return r->hip_fail(e, "image readback");
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_read_energy(wf_renderer* r, float* out)
{
    if (!r || !out) return WF_ERROR_INVALID_ARGUMENT;
    return read_image(r, r->d_image, out, 16ull * r->width * r->height);
}

extern "C" WF_API wf_status wf_renderer_read_accumulator(wf_renderer* r, float* out)
{
    if (!r || !out) return WF_ERROR_INVALID_ARGUMENT;
    return read_image(r, r->d_accum, out, 16ull * r->width * r->height);
}

extern "C" WF_API wf_status wf_renderer_read_final_color(wf_renderer* r, uint8_t* out)
{
    if (!r || !out) return WF_ERROR_INVALID_ARGUMENT;
    return read_image(r, r->d_final, out, 4ull * r->width * r->height);
}
