// Frame / scene lifecycle (renderer.cpp:543-630, 1051-1055), the material
// and emitter tables the shade and connect stages read, and the frame loop
// (renderer.cpp:742-939) over one or many scene slots.
#include "renderer_impl.h"
#include "mega.h"

#include <algorithm>
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

static wf_status open_slot(wf_renderer* r, const wf_camera* cam, uint32_t w, uint32_t h, uint32_t keep_accum)
{
    SceneSlot sl;
    sl.camera = *cam;
    sl.width = w;
    sl.height = h;
    sl.accum_count = keep_accum;
    make_view(*cam, w, h, sl.view);
    r->slots.push_back(std::move(sl));
    r->cur_slot = (uint32_t)r->slots.size() - 1;
    r->tlas_valid = false;
    return WF_OK;
}

// begin_scene (renderer.cpp:599-630): the scene becomes slot 0; the
// accumulator is kept only while the view matrix is unchanged.
extern "C" WF_API wf_status wf_renderer_begin_scene(wf_renderer* r, const wf_camera* cam, wf_texture_handle env)
{
    if (!r || !cam) return WF_ERROR_INVALID_ARGUMENT;
    if (r->scene_open) return r->fail(WF_ERROR_INVALID_STATE, "begin_scene called twice");
    if (!(cam->vfov_deg > 0.0f && cam->vfov_deg < 180.0f)) return r->fail(WF_ERROR_INVALID_ARGUMENT, "vfov_deg out of range");
    uint32_t keep = 0;
    if (r->slots.size() == 1 && std::memcmp(cam->view_matrix, r->slots[0].camera.view_matrix, 64) == 0 &&
        r->slots[0].width == r->width && r->slots[0].height == r->height)
        keep = r->slots[0].accum_count;
    r->slots.clear();
    r->env_texture = env;
    r->scene_open = true;
    return open_slot(r, cam, r->width, r->height, keep);
}

// end_scene (renderer.cpp:1051-1055): the instances stay until the next begin_scene
extern "C" WF_API wf_status wf_renderer_end_scene(wf_renderer* r)
{
    if (!r) return WF_ERROR_INVALID_ARGUMENT;
    if (!r->scene_open) return r->fail(WF_ERROR_INVALID_STATE, "end_scene without begin_scene");
    r->scene_open = false;
    return WF_OK;
}

// ---- batches: many scenes of their own camera and resolution ---------------

extern "C" WF_API wf_status wf_renderer_begin_batch(wf_renderer* r)
{
    if (!r) return WF_ERROR_INVALID_ARGUMENT;
    if (r->scene_open) return r->fail(WF_ERROR_INVALID_STATE, "begin_batch inside an open scene");
    r->slots.clear();
    r->env_texture = WF_INVALID_HANDLE;
    r->tlas_valid = false;
    r->scene_open = true;
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_batch_add_scene(wf_renderer* r, const wf_camera* cam, uint32_t width, uint32_t height,
                                                        uint32_t* out_scene)
{
    if (!r || !cam || width == 0 || height == 0) return WF_ERROR_INVALID_ARGUMENT;
    if (!r->scene_open) return r->fail(WF_ERROR_INVALID_STATE, "batch_add_scene without begin_batch");
    if (!(cam->vfov_deg > 0.0f && cam->vfov_deg < 180.0f)) return r->fail(WF_ERROR_INVALID_ARGUMENT, "vfov_deg out of range");
    if ((uint64_t)width * height > (1ull << 26)) return r->fail(WF_ERROR_CAPACITY, "scene resolution too large");
    if (r->slots.size() >= 4096) return r->fail(WF_ERROR_CAPACITY, "at most 4096 scenes per batch");
    wf_status s = open_slot(r, cam, width, height, 0);
    if (s == WF_OK && out_scene) *out_scene = r->cur_slot;
    return s;
}

extern "C" WF_API wf_status wf_renderer_get_scene_count(const wf_renderer* r, uint32_t* out)
{
    if (!r || !out) return WF_ERROR_INVALID_ARGUMENT;
    *out = (uint32_t)r->slots.size();
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_get_view(const wf_renderer* r, wf_view* out)
{
    if (!r || !out) return WF_ERROR_INVALID_ARGUMENT;
    if (r->slots.empty()) return WF_ERROR_INVALID_STATE;
    *out = r->slots[0].view;
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_get_frame_stats(const wf_renderer* r, wf_frame_stats* out)
{
    if (!r || !out) return WF_ERROR_INVALID_ARGUMENT;
    *out = r->stats;
    return WF_OK;
}

void wf_renderer::drop_graph()
{
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    graph_exec = nullptr;
    graph_key.clear();
    forward_valid = false;
}

void wf_renderer::release_frame_buffers()
{
    drop_graph();
    for (auto& b : d_rays) b.release();
    for (auto& b : d_paths) b.release();
    for (auto& b : d_roots) b.release();
    d_hits.release(); d_thr.release(); d_path_energy.release(); d_path_scene.release(); d_counts.release();
    d_shadow_rays.release(); d_shadow_path.release(); d_shadow_root.release(); d_shadow_contrib.release(); d_shadow_visible.release();
    d_image.release(); d_accum.release(); d_final.release();
    d_scenes.release(); d_dyn.release();
    frame_capacity = 0;
    image_pixels = 0;
    image_valid = false;
}

// Texture table and, per slot, the emitter list for next-event estimation:
// every triangle of an instance whose emissive factor is non-zero, with its
// world area (local triangle transformed by local_to_world), as a CDF.
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

    std::vector<uint32_t> rec;   // (instance, primitive, cdf bits, 0)
    for (auto& sl : slots) {
        std::vector<uint32_t> ids;
        std::vector<double> area;
        for (size_t i = 0; i < sl.instances.size(); ++i) {
            const InstanceRec& ir = sl.instances[i];
            const float em = ir.material.emissive_strength;
            const float* ef = ir.material.emissive_factor;
            if (!(em * ef[0] > 0.0f || em * ef[1] > 0.0f || em * ef[2] > 0.0f)) continue;
            const Mesh& m = meshes[ir.mesh];
            for (uint32_t t = 0; t < m.tri_count; ++t) {
                float w[3][4];
                for (int v = 0; v < 3; ++v) {
                    const float p[4] = {m.pos9[9ull * t + 3 * v], m.pos9[9ull * t + 3 * v + 1], m.pos9[9ull * t + 3 * v + 2], 1.0f};
                    mat4_mul_vec4(ir.transform, p, w[v]);
                }
                const double e1[3] = {(double)w[1][0] - w[0][0], (double)w[1][1] - w[0][1], (double)w[1][2] - w[0][2]};
                const double e2[3] = {(double)w[2][0] - w[0][0], (double)w[2][1] - w[0][1], (double)w[2][2] - w[0][2]};
                const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
                const double a = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
                if (!(a > 0.0)) continue;
                ids.push_back(sl.inst_base + (uint32_t)i);
                ids.push_back(t);
                area.push_back(a);
            }
        }
        double total = 0.0;
        for (double a : area) total += a;
        sl.em_off = (uint32_t)(rec.size() / 4);
        sl.em_count = (uint32_t)area.size();
        sl.em_area = (float)total;
        double acc = 0.0;
        for (size_t k = 0; k < area.size(); ++k) {
            acc += area[k];
            float c = (float)(acc / total);
            if (k + 1 == area.size()) c = 1.0f;
            uint32_t cb;
            std::memcpy(&cb, &c, 4);
            rec.insert(rec.end(), {ids[2 * k], ids[2 * k + 1], cb, 0u});
        }
    }
    if (rec.empty()) rec.assign(4, 0u);
    e = d_emitters.ensure(4 * rec.size());
    if (e == hipSuccess) e = hipMemcpy(d_emitters.p, rec.data(), 4 * rec.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "emitter table upload");
    return WF_OK;
}

// Queues sized for every path of the batch (the reference sizes them to
// W*H, renderer.cpp:432-469); the accumulator and output images per pixel.
wf_status wf_renderer::ensure_frame_buffers(size_t n_paths, size_t n_pixels)
{
    hipError_t e = hipSuccess;
    bool moved = false;
    auto need = [&](DevBuf& b, size_t bytes) {
        bytes = std::max<size_t>(bytes, 64);
        if (e != hipSuccess || (b.p && b.bytes >= bytes)) return;
        moved = true;
        e = b.ensure(bytes);
    };
    for (auto& b : d_rays) need(b, sizeof(wf_ray) * n_paths);
    for (auto& b : d_paths) need(b, 4 * n_paths);
    for (auto& b : d_roots) need(b, 4 * n_paths);
    need(d_hits, sizeof(wf_hit) * n_paths);
    need(d_thr, 16 * n_paths);
    need(d_path_energy, 16 * n_paths);
    need(d_path_scene, 4 * n_paths);
    need(d_counts, 4 * kCountSlots);
    need(d_shadow_rays, sizeof(wf_ray) * n_paths);
    need(d_shadow_path, 4 * n_paths);
    need(d_shadow_root, 4 * n_paths);
    need(d_shadow_contrib, 16 * n_paths);
    need(d_shadow_visible, 4 * n_paths);
    const void* accum_before = d_accum.p;
    const bool fresh_images = n_pixels != image_pixels;
    need(d_image, 16 * n_pixels);
    need(d_accum, 16 * n_pixels);
    need(d_final, 4 * n_pixels);
    need(d_scenes, sizeof(GpuScene) * slots.size());
    need(d_dyn, sizeof(FrameDyn));
    if (e != hipSuccess) return hip_fail(e, "frame buffer allocation");
    if (moved) drop_graph();
    if (fresh_images || d_accum.p != accum_before) {
        // the accumulator starts zeroed, as the reference's render target
        // does (renderer.cpp:484-487): a first frame's weight-1 blend,
        // accum * 0 + energy, is then exactly the frame
        for (auto& sl : slots) sl.accum_count = 0;
        const size_t words = 4 * n_pixels, chunk = size_t(1) << 28;
        for (size_t off = 0; off < words && e == hipSuccess; off += chunk)
            e = launch_fill_u32((uint32_t*)d_accum.p + off, 0u, (uint32_t)std::min(chunk, words - off), stream);
        if (e != hipSuccess) return hip_fail(e, "accumulator clear");
    }
    image_pixels = n_pixels;
    frame_capacity = n_paths;
    return WF_OK;
}

// The launches of one frame: clear, generate, then for every depth d <=
// max_bounces: extend, shade, connect (shadow extend + accumulate); then
// resolve + post-process.  With `timed`, events separate the stages.
wf_status wf_renderer::record_frame(const FrameArgs& A, uint32_t grid, bool timed, int* n_events)
{
    hipError_t e;
    int ei = 0;
    auto mark = [&]() -> hipError_t { return timed ? hipEventRecord(ev[ei++], stream) : hipSuccess; };
    wf_status s;
    if ((s = reset_overflows()) != WF_OK) return s;
    if ((e = mark()) != hipSuccess) return hip_fail(e, "event");
    if (!settings.use_wavefront_pathtracing) {
        // megakernel frame: every path from camera to termination in one launch
        MegaLaunch M{};
        M.scene.tnodes = (const wf_tlas_node*)d_tnodes.p;
        M.scene.tpairs = (const float4*)d_tpairs.p;
        M.scene.n_tpairs = n_tpairs;
        M.scene.inst = (const GpuInstance*)d_inst.p;
        M.scene.tquads = (const float4*)d_tquads.p;
        M.spill = (uint2*)d_spill.p;
        M.spill_per_thread = spill_per_thread;
        M.stack_cap = ext_cfg.stack_limit ? std::min(ext_cfg.stack_limit, kStackTotal) : kStackTotal;
        M.work = (uint32_t*)d_work.p;
        M.visits = ext_cfg.visits;
        M.refill_min = ext_cfg.refill_min;
        M.grid = mega_grid;
        if ((e = launch_megakernel(A, M, ext_cfg, stream)) != hipSuccess) return hip_fail(e, "megakernel launch");
        if ((e = mark()) != hipSuccess) return hip_fail(e, "event");
        if ((e = launch_resolve_post(A, grid, stream)) != hipSuccess) return hip_fail(e, "post-process launch");
        if ((e = mark()) != hipSuccess) return hip_fail(e, "event");
        *n_events = ei;
        return WF_OK;
    }
    if ((e = launch_clear(A, grid, stream)) != hipSuccess) return hip_fail(e, "clear launch");
    if ((e = mark()) != hipSuccess) return hip_fail(e, "event");
    if ((e = launch_generate(A, grid, stream)) != hipSuccess) return hip_fail(e, "generate launch");
    if ((e = mark()) != hipSuccess) return hip_fail(e, "event");
    // The shadow rays of bounce d-1 are traced in the launch that extends
    // bounce d (one launch per bounce instead of two half-full ones), and
    // their contributions are added before bounce d is shaded: every path
    // sees the same operations in the same order as with separate launches.
    // separate_connect keeps the reference's one launch per stage.
    const bool batched = slots.size() > 1;
    // (the fused kernels exist with the TLAS in global memory, and for the
    // 4-wide layout with the popped-interior skip)
    const bool fuse = settings.next_event_estimation && !settings.separate_connect && (ext_cfg.skip_popped || ext_cfg.layout == 0) &&
                      ext_grid_tlds == 0;
    for (uint32_t d = 0; d <= settings.max_bounces; ++d) {
        const bool with_shadow = fuse && d > 0;
        OcclusionQueue q{A.shadow_rays, A.counts + kShadowCountBase + (d > 0 ? d - 1 : 0), A.shadow_visible,
                         batched ? A.shadow_root : nullptr};
        s = run_extend(A.rays[d & 1u], A.counts + d, A.n_paths, A.hits, nullptr, nullptr, false,
                       batched ? A.root_of[d & 1u] : nullptr, with_shadow ? &q : nullptr);
        if (s != WF_OK) return s;
        if (with_shadow && (e = launch_connect_accumulate(A, d - 1, grid, stream)) != hipSuccess)
            return hip_fail(e, "connect launch");
        if ((e = mark()) != hipSuccess) return hip_fail(e, "event");
        if ((e = launch_shade(A, d, grid, stream)) != hipSuccess) return hip_fail(e, "shade launch");
        if ((e = mark()) != hipSuccess) return hip_fail(e, "event");
        if (settings.next_event_estimation && !fuse && d < settings.max_bounces) {
            s = run_extend(A.shadow_rays, A.counts + kShadowCountBase + d, A.n_paths, nullptr, A.shadow_visible, nullptr, true,
                           batched ? A.shadow_root : nullptr);
            if (s != WF_OK) return s;
            if ((e = launch_connect_accumulate(A, d, grid, stream)) != hipSuccess) return hip_fail(e, "connect launch");
        }
        if ((e = mark()) != hipSuccess) return hip_fail(e, "event");
    }
    if ((e = launch_resolve_post(A, grid, stream)) != hipSuccess) return hip_fail(e, "post-process launch");
    if ((e = mark()) != hipSuccess) return hip_fail(e, "event");
    *n_events = ei;
    return WF_OK;
}

// renderer::render (renderer.cpp:632-1049), wavefront branch, for the slots
// [first, first + count).  Path ids are numbered over all slots, so a scene
// renders the same image whether it is drawn alone or inside a batch.
wf_status wf_renderer::render_range(uint32_t first, uint32_t count)
{
    if (slots.empty()) return fail(WF_ERROR_INVALID_STATE, "render called before begin_scene");
    if (count == 0 || first >= slots.size() || count > slots.size() - first)
        return fail(WF_ERROR_INVALID_ARGUMENT, "scene range [%u, %u) outside [0, %zu)", first, first + count, slots.size());
    wf_status s = upload_scene();
    if (s != WF_OK) return s;

    uint64_t path_total = 0, pix_total = 0;
    for (auto& sl : slots) {
        sl.path_off = (uint32_t)path_total;
        sl.pix_off = (uint32_t)pix_total;
        const uint64_t np = (uint64_t)sl.width * sl.height;
        path_total += np * settings.spp;
        pix_total += np;
        if (path_total > (1ull << 30)) return fail(WF_ERROR_CAPACITY, "the scenes' paths (pixels x spp) exceed 2^30");
    }
    if ((s = ensure_frame_buffers(path_total, pix_total)) != WF_OK) return s;

    // per-frame values: scene table (accumulation weights) and the seed
    h_scenes.assign(sizeof(GpuScene) * slots.size(), 0);
    GpuScene* gs = reinterpret_cast<GpuScene*>(h_scenes.data());
    for (size_t k = 0; k < slots.size(); ++k) {
        SceneSlot& sl = slots[k];
        if (k >= first && k < first + count && sl.accum_count == 0) sl.accum_count = 1;   // reset_color_accumulator
        gs[k].view = sl.view;
        gs[k].width = sl.width;
        gs[k].height = sl.height;
        gs[k].n_pix = sl.width * sl.height;
        gs[k].pix_off = sl.pix_off;
        gs[k].path_off = sl.path_off;
        gs[k].n_paths = gs[k].n_pix * settings.spp;
        gs[k].tlas_root = sl.tlas_root;
        gs[k].em_off = sl.em_off;
        gs[k].em_count = sl.em_count;
        gs[k].em_area = sl.em_area;
        gs[k].sample_count = std::max<uint32_t>(sl.accum_count, 1);
    }
    // renderer.cpp:739: a new frame seed per frame unless the caller fixes it
    uint32_t seed = settings.frame_seed;
    if (seed == 0) {
        xorshift_seed ^= xorshift_seed << 13;
        xorshift_seed ^= xorshift_seed >> 17;
        xorshift_seed ^= xorshift_seed << 5;
        seed = xorshift_seed;
    }
    h_dyn[0] = seed;
    hipError_t e = hipMemcpyAsync(d_scenes.p, h_scenes.data(), h_scenes.size(), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_dyn.p, h_dyn, sizeof(h_dyn), hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return hip_fail(e, "frame constants upload");

    const SceneSlot& f = slots[first];
    const SceneSlot& l = slots[first + count - 1];
    FrameArgs A{};
    A.scenes = (const GpuScene*)d_scenes.p;
    A.first_scene = first;
    A.n_scenes = count;
    A.path_base = f.path_off;
    A.n_paths = (uint32_t)(l.path_off + (uint64_t)l.width * l.height * settings.spp - f.path_off);
    A.pix_base = f.pix_off;
    A.n_pix = l.pix_off + l.width * l.height - f.pix_off;
    A.spp = settings.spp;
    A.max_bounces = settings.max_bounces;
    A.rr_start = settings.rr_start_bounce;
    A.nee = settings.next_event_estimation;
    A.uniform_diffuse = settings.cosine_weighted_diffuse ? 0u : 1u;
    A.view_mode = settings.render_view_mode;
    A.rr_survival = settings.rr_survival;
    for (int k = 0; k < 3; ++k) A.env[k] = settings.hdr_env_strength;   // default 1x1 white environment
    // begin_scene's environment texture; a missing one falls back to the
    // white default as the reference does (renderer.cpp:605-610)
    A.env_tex = (env_texture != WF_INVALID_HANDLE && find_texture(env_texture)) ? (int32_t)(env_texture & 0xFFFFFFFFu) : -1;
    A.env_strength = settings.hdr_env_strength;
    A.accumulate = settings.accumulate;
    A.dyn = (const FrameDyn*)d_dyn.p;
    A.thr = (float4*)d_thr.p;
    A.energy = (float4*)d_path_energy.p;
    A.path_scene = (uint32_t*)d_path_scene.p;
    for (int k = 0; k < 2; ++k) {
        A.rays[k] = (wf_ray*)d_rays[k].p;
        A.path_of[k] = (uint32_t*)d_paths[k].p;
        A.root_of[k] = (uint32_t*)d_roots[k].p;
    }
    A.hits = (wf_hit*)d_hits.p;
    A.counts = (uint32_t*)d_counts.p;
    A.shadow_rays = (wf_ray*)d_shadow_rays.p;
    A.shadow_path = (uint32_t*)d_shadow_path.p;
    A.shadow_root = (uint32_t*)d_shadow_root.p;
    A.shadow_contrib = (float4*)d_shadow_contrib.p;
    A.shadow_visible = (uint32_t*)d_shadow_visible.p;
    A.shade_inst = (const GpuShadeInstance*)d_shade_inst.p;
    A.textures = (const GpuTexture*)d_textures.p;
    A.emitters = (const uint4*)d_emitters.p;
    A.image = (float4*)d_image.p;
    A.accum = (float4*)d_accum.p;
    A.final_rgba8 = (uint32_t*)d_final.p;
    const uint32_t grid = (uint32_t)cu_count * 8u;

    wf_frame_stats st{};
    int n_ev = 0;
    if (settings.use_hip_graph) {
        // the frame as one captured graph; everything that changes between
        // frames is read from device memory, so the graph is reused while the
        // buffers, scene range, settings and traversal kernel stay the same
        auto bits = [](float f) { uint32_t u; std::memcpy(&u, &f, 4); return (uint64_t)u; };
        const std::vector<uint64_t> key = {
            first, count, settings.use_wavefront_pathtracing, mega_grid, settings.max_bounces, settings.spp, settings.next_event_estimation, settings.rr_start_bounce,
            bits(settings.rr_survival), bits(settings.hdr_env_strength),
            settings.accumulate, ext_cfg.block, ext_cfg.lds_stack, ext_cfg.skip_popped, ext_grid,
            (uint64_t)(uintptr_t)d_tnodes.p, (uint64_t)(uintptr_t)d_tpairs.p, ext_cfg.refill_min, ext_cfg.tlas_lds, ext_cfg.visits, ext_grid_tlds, n_tpairs, (uint64_t)(uintptr_t)d_inst.p, (uint64_t)(uintptr_t)d_shade_inst.p,
            (uint64_t)(uintptr_t)d_emitters.p, (uint64_t)(uintptr_t)d_textures.p, (uint64_t)(uintptr_t)d_spill.p,
            (uint64_t)(uintptr_t)d_rays[0].p, (uint64_t)(uintptr_t)d_image.p, slots.size(), ext_cfg.layout, settings.separate_connect, (uint64_t)(uint32_t)A.env_tex,
            (uint64_t)(uintptr_t)d_tquads.p, A.uniform_diffuse, A.view_mode, ext_cfg.big};
        if (!graph_exec || key != graph_key) {
            drop_graph();
            hipGraph_t g = nullptr;
            if ((e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal)) != hipSuccess) return hip_fail(e, "graph capture");
            s = record_frame(A, grid, false, &n_ev);
            e = hipStreamEndCapture(stream, &g);
            if (s != WF_OK) { if (g) (void)hipGraphDestroy(g); return s; }
            if (e != hipSuccess) return hip_fail(e, "graph capture");
            e = hipGraphInstantiate(&graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) { graph_exec = nullptr; return hip_fail(e, "graph instantiate"); }
            graph_key = key;
        }
        if ((e = hipEventRecord(ev[0], stream)) != hipSuccess) return hip_fail(e, "event");
        if ((e = hipGraphLaunch(graph_exec, stream)) != hipSuccess) return hip_fail(e, "graph launch");
        if ((e = hipEventRecord(ev[1], stream)) != hipSuccess) return hip_fail(e, "event");
        n_ev = 2;
    } else {
        if ((s = record_frame(A, grid, true, &n_ev)) != WF_OK) return s;
    }

    uint32_t counts[kCountSlots];
    e = hipMemcpyAsync(counts, d_counts.p, sizeof(counts), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "frame");
    auto ms = [&](int a, int b) { float t = 0.0f; (void)hipEventElapsedTime(&t, ev[a], ev[b]); return t; };
    if (!settings.use_hip_graph && !settings.use_wavefront_pathtracing) {
        st.ms_post = ms(1, 2);   // the megakernel frame has no separate stages
    } else if (!settings.use_hip_graph) {
        st.ms_clear = ms(0, 1);
        st.ms_generate = ms(1, 2);
        for (uint32_t d = 0; d <= settings.max_bounces; ++d) {
            const int b = 2 + 3 * (int)d;
            st.ms_extend += ms(b, b + 1);
            st.ms_shade += ms(b + 1, b + 2);
            st.ms_connect += ms(b + 2, b + 3);
        }
        st.ms_post = ms(n_ev - 2, n_ev - 1);
    }
    st.ms_total = ms(0, n_ev - 1);
    for (uint32_t d = 0; d <= settings.max_bounces; ++d) {
        st.rays_extended += counts[d];
        st.rays_per_bounce[d] = counts[d];
        if (d < kMaxBounces) {
            st.shadow_rays += counts[kShadowCountBase + d];
            st.shadow_rays_per_bounce[d] = counts[kShadowCountBase + d];
        }
    }
    uint32_t ov = 0;
    if ((s = read_overflows(&ov)) != WF_OK) return s;
    st.stack_overflows = ov;
    stats = st;
    image_valid = true;
    last_args = A;
    forward_valid = true;
    // renderer.cpp:593-596: the sample count grows only while accumulating
    if (settings.accumulate)
        for (uint32_t k = first; k < first + count; ++k) slots[k].accum_count++;
    if (ov) return fail(WF_ERROR_STACK_OVERFLOW, "%u rays overflowed the traversal stack this frame", ov);
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_render(wf_renderer* r)
{
    if (!r) return WF_ERROR_INVALID_ARGUMENT;
    return r->render_range(0, (uint32_t)r->slots.size());
}

extern "C" WF_API wf_status wf_renderer_render_scenes(wf_renderer* r, uint32_t first, uint32_t count)
{
    if (!r) return WF_ERROR_INVALID_ARGUMENT;
    return r->render_range(first, count);
}

static wf_status read_image(wf_renderer* r, const DevBuf& b, uint32_t slot, size_t texel, void* out)
{
    if (!r->image_valid) return r->fail(WF_ERROR_INVALID_STATE, "no frame rendered");
    if (slot >= r->slots.size()) return r->fail(WF_ERROR_INVALID_ARGUMENT, "scene %u out of range", slot);
    const SceneSlot& sl = r->slots[slot];
    hipError_t e = hipMemcpyAsync(out, (const uint8_t*)b.p + texel * sl.pix_off, texel * sl.width * sl.height,
                                  hipMemcpyDeviceToHost, r->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) return r->hip_fail(e, "image readback");
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_read_energy(wf_renderer* r, float* out)
{
    if (!r || !out) return WF_ERROR_INVALID_ARGUMENT;
    return read_image(r, r->d_image, 0, 16, out);
}

extern "C" WF_API wf_status wf_renderer_read_accumulator(wf_renderer* r, float* out)
{
    if (!r || !out) return WF_ERROR_INVALID_ARGUMENT;
    return read_image(r, r->d_accum, 0, 16, out);
}

extern "C" WF_API wf_status wf_renderer_read_final_color(wf_renderer* r, uint8_t* out)
{
    if (!r || !out) return WF_ERROR_INVALID_ARGUMENT;
    return read_image(r, r->d_final, 0, 4, out);
}

extern "C" WF_API wf_status wf_renderer_read_scene_energy(wf_renderer* r, uint32_t scene, float* out)
{
    if (!r || !out) return WF_ERROR_INVALID_ARGUMENT;
    return read_image(r, r->d_image, scene, 16, out);
}
