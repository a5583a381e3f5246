// Host runtime: the renderer object behind the C-ABI in include/wf.h.
//
// Mirrors renderer::init / create_render_mesh / submit_render_mesh /
// begin_scene / render / end_scene (renderer.cpp:222-1350) with HIP device
// memory in place of D3D12 resources and an explicit handle in place of the
// g_renderer singleton.  Software ray tracing is the only traversal mode.
#include "renderer_impl.h"
#include "mega.h"

#include <algorithm>
#include <cstdio>
#include <cstring>

using namespace wf;

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

wf_status wf_renderer::fail(wf_status s, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    err = buf;
    return s;
}

wf_status wf_renderer::hip_fail(hipError_t e, const char* what)
{
    const wf_status s = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? WF_ERROR_OUT_OF_MEMORY : WF_ERROR_DEVICE;
    return fail(s, "%s: %s", what, hipGetErrorString(e));
}

hipError_t DevBuf::ensure(size_t n)
{
    if (n <= bytes && p) return hipSuccess;
    release();
    if (n == 0) return hipSuccess;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) { p = nullptr; bytes = 0; return e; }
    bytes = n;
    return hipSuccess;
}

void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

static inline uint64_t make_handle(uint32_t index, uint32_t version) { return ((uint64_t)version << 32) | index; }
static inline uint32_t handle_index(uint64_t h) { return (uint32_t)(h & 0xFFFFFFFFu); }
static inline uint32_t handle_version(uint64_t h) { return (uint32_t)(h >> 32); }

Mesh* wf_renderer::find_mesh(uint64_t h)
{
    if (h == WF_INVALID_HANDLE) return nullptr;
    const uint32_t i = handle_index(h);
    if (i >= meshes.size() || !meshes[i].alive || meshes[i].version != handle_version(h)) return nullptr;
    return &meshes[i];
}

Texture* wf_renderer::find_texture(uint64_t h)
{
    if (h == WF_INVALID_HANDLE) return nullptr;
    const uint32_t i = handle_index(h);
    if (i >= textures.size() || !textures[i].alive || textures[i].version != handle_version(h)) return nullptr;
    return &textures[i];
}

// get_default_render_settings (renderer.cpp:32-45) + this library's additions
static wf_render_settings default_settings()
{
    wf_render_settings s{};
    s.use_wavefront_pathtracing = 1;
    s.use_software_rt = 1;
    s.render_view_mode = 0;
    s.max_bounces = 3;
    s.cosine_weighted_diffuse = 1;
    s.accumulate = 1;
    s.hdr_env_strength = 1.0f;
    s.spp = 1;
    s.rr_start_bounce = 3;
    s.rr_survival = 0.75f;
    s.next_event_estimation = 0;
    s.frame_seed = 0;
    s.use_hip_graph = 0;
    s.separate_connect = 0;
    return s;
}

// Grid of the LDS-TLAS variant for the current TLAS (0 = use the global
// variant): its dynamic LDS grows with the TLAS, so its occupancy does too.
wf_status wf_renderer::update_tlds_grid()
{
    ext_grid_tlds = 0;
    if (!ext_cfg.tlas_lds || n_tpairs == 0 || n_tpairs > kTlasLdsPairs) return WF_OK;
    // stay within 64 KB of LDS per workgroup (stack + TLAS copy)
    if ((size_t)ext_cfg.lds_stack * ext_cfg.block * 8u + (size_t)n_tpairs * 64u > 65536u) return WF_OK;
    int per_cu = 0;
    const hipError_t e = extend_occupancy(ext_cfg, true, (size_t)n_tpairs * 64u, &per_cu);
    if (e != hipSuccess) return hip_fail(e, "occupancy query");
    if (per_cu >= 1) ext_grid_tlds = std::min((uint32_t)per_cu * (uint32_t)cu_count, ext_grid);
    return WF_OK;
}

wf_status wf_renderer::ensure_trace_scratch()
{
    // persistent grid: as many workgroups as stay resident on every CU
    int per_cu = 0;
    hipError_t e = extend_occupancy(ext_cfg, false, 0, &per_cu);
    if (e != hipSuccess) return hip_fail(e, "occupancy query");
    if (per_cu < 1) per_cu = 1;
    ext_grid = (uint32_t)per_cu * (uint32_t)cu_count;
    wf_status s = update_tlds_grid();
    if (s != WF_OK) return s;
    int per_cu_m = 0;
    if ((e = mega_occupancy(ext_cfg, &per_cu_m)) != hipSuccess) return hip_fail(e, "occupancy query");
    mega_grid = (uint32_t)std::max(per_cu_m, 1) * (uint32_t)cu_count;
    spill_per_thread = ext_cfg.lds_stack > 0 ? (kStackTotal > ext_cfg.lds_stack ? kStackTotal - ext_cfg.lds_stack : 0) : 0;
    const size_t threads = (size_t)std::max(ext_grid, mega_grid) * ext_cfg.block;
    if ((e = d_spill.ensure(std::max<size_t>(threads * spill_per_thread * sizeof(uint2), 16))) != hipSuccess)
        return hip_fail(e, "spill stack allocation");
    if ((e = d_work.ensure(kWorkWords * sizeof(uint32_t))) != hipSuccess) return hip_fail(e, "work counter allocation");
    return WF_OK;
}

// ---------------------------------------------------------------------------
// lifecycle
// ---------------------------------------------------------------------------

extern "C" WF_API wf_status wf_renderer_init(const wf_init_params* params, wf_renderer** out)
{
    if (!params || !out) return WF_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (params->render_width == 0 || params->render_height == 0 ||
        (uint64_t)params->render_width * params->render_height > (1ull << 28))
        return WF_ERROR_INVALID_ARGUMENT;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return WF_ERROR_DEVICE;
    wf_renderer* r = new (std::nothrow) wf_renderer();
    if (!r) return WF_ERROR_OUT_OF_MEMORY;
    r->width = params->render_width;
    r->height = params->render_height;
    r->settings = default_settings();
    r->build_opts = wf_build_options{8, 0};   // renderer.cpp:95-96
    hipError_t e = hipGetDevice(&r->device);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, r->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete r; return WF_ERROR_DEVICE; }
    r->cu_count = prop.multiProcessorCount;
    for (int i = 0; i < kNumStageEvents; ++i) {
        if (hipEventCreate(&r->ev[i]) != hipSuccess) { wf_renderer_exit(r); return WF_ERROR_DEVICE; }
    }
    if (r->ensure_trace_scratch() != WF_OK) { wf_renderer_exit(r); return WF_ERROR_DEVICE; }
    *out = r;
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_exit(wf_renderer* r)
{
    if (!r) return WF_ERROR_INVALID_ARGUMENT;
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    for (auto& m : r->meshes) { m.nodes.release(); m.tris.release(); m.shade.release(); m.quads.release(); }
    r->d_tquads.release();
    r->build_scratch.release();
    for (wf::DevBuf* d : {&r->tbuild_inst, &r->tbuild_nodes, &r->tbuild_list, &r->tbuild_ids, &r->tbuild_status}) d->release();
    r->build_pos.release();
    r->build_idx.release();
    for (auto& t : r->textures) t.texels.release();
    r->release_frame_buffers();
    r->d_tnodes.release(); r->d_tpairs.release(); r->d_inst.release(); r->d_shade_inst.release(); r->d_textures.release();
    r->d_emitters.release(); r->d_scenes.release(); r->d_dyn.release();
    r->d_lrem.release(); r->d_delta.release(); r->d_shadow_unit.release(); r->d_shadow_light.release();
    r->d_adj_dl.release(); r->d_adj_tex.release(); r->d_adj_em.release();
    r->d_work.release(); r->d_check.release(); r->d_spill.release(); r->d_tmp_rays.release(); r->d_tmp_hits.release(); r->d_tmp_aux.release();
    r->d_order.release(); r->d_order_scratch.release();
    for (int i = 0; i < kNumStageEvents; ++i) if (r->ev[i]) (void)hipEventDestroy(r->ev[i]);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
    return WF_OK;
}

extern "C" WF_API const char* wf_renderer_last_error(const wf_renderer* r)
{
    return r ? r->err.c_str() : "null renderer";
}

extern "C" WF_API wf_status wf_renderer_get_stream(const wf_renderer* r, void** out)
{
    if (!r || !out) return WF_ERROR_INVALID_ARGUMENT;
    *out = (void*)r->stream;
    return WF_OK;
}

// ---------------------------------------------------------------------------
// settings / camera
// ---------------------------------------------------------------------------

extern "C" WF_API wf_status wf_renderer_get_settings(const wf_renderer* r, wf_render_settings* out)
{
    if (!r || !out) return WF_ERROR_INVALID_ARGUMENT;
    *out = r->settings;
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_set_settings(wf_renderer* r, const wf_render_settings* s)
{
    if (!r || !s) return WF_ERROR_INVALID_ARGUMENT;
    if (s->max_bounces > kMaxBounces) return r->fail(WF_ERROR_INVALID_ARGUMENT, "max_bounces %u > %u", s->max_bounces, kMaxBounces);
    if (s->spp == 0 || s->spp > 4096) return r->fail(WF_ERROR_INVALID_ARGUMENT, "spp must be in [1, 4096]");
    if (!(s->rr_survival > 0.0f && s->rr_survival <= 1.0f)) return r->fail(WF_ERROR_INVALID_ARGUMENT, "rr_survival must be in (0, 1]");
    if (s->use_software_rt == 0) return r->fail(WF_ERROR_INVALID_ARGUMENT, "only software ray tracing is available");
    if (s->render_view_mode >= WF_RENDER_VIEW_MODE_COUNT)
        return r->fail(WF_ERROR_INVALID_ARGUMENT, "render_view_mode %u is not a view mode", s->render_view_mode);
    // renderer.cpp:1099-1151 resets the accumulator when the view mode, max
    // bounces, accumulate, cosine-weighted sampling or the environment strength
    // change; spp and next-event estimation have no reference counterpart and
    // change the image too
    const bool reset = s->max_bounces != r->settings.max_bounces || s->spp != r->settings.spp ||
                       s->next_event_estimation != r->settings.next_event_estimation ||
                       s->render_view_mode != r->settings.render_view_mode || s->accumulate != r->settings.accumulate ||
                       s->cosine_weighted_diffuse != r->settings.cosine_weighted_diffuse ||
                       std::memcmp(&s->hdr_env_strength, &r->settings.hdr_env_strength, sizeof(float)) != 0;
    if (std::memcmp(s, &r->settings, sizeof(*s)) != 0) r->forward_valid = false;
    r->settings = *s;
    if (reset)
        for (auto& sl : r->slots) sl.accum_count = 0;
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_set_traversal_config(wf_renderer* r, const wf_traversal_config* c)
{
    if (!r || !c) return WF_ERROR_INVALID_ARGUMENT;
    if (c->workgroup_size != 64 && c->workgroup_size != 128 && c->workgroup_size != 256)
        return r->fail(WF_ERROR_INVALID_ARGUMENT, "workgroup_size must be 64, 128 or 256");
    if (c->lds_stack_entries != 0 && c->lds_stack_entries != 4 && c->lds_stack_entries != 8 && c->lds_stack_entries != 12 &&
        c->lds_stack_entries != 16)
        return r->fail(WF_ERROR_INVALID_ARGUMENT, "lds_stack_entries must be 0, 4, 8, 12 or 16");
    if (c->bvh_layout > 1) return r->fail(WF_ERROR_INVALID_ARGUMENT, "bvh_layout must be 0 (2-wide) or 1 (4-wide)");
    if (c->bvh_layout == 1 && (!c->skip_popped_interior || c->tlas_in_lds))
        return r->fail(WF_ERROR_INVALID_ARGUMENT, "the 4-wide layout needs skip_popped_interior = 1 and tlas_in_lds = 0");
    if (c->refill_min > 64) return r->fail(WF_ERROR_INVALID_ARGUMENT, "refill_min must be in [0, 64]");
    if (c->interior_visits > 8) return r->fail(WF_ERROR_INVALID_ARGUMENT, "interior_visits must be in [0, 8]");
    if (c->big_claims > 2) return r->fail(WF_ERROR_INVALID_ARGUMENT, "big_claims must be 0 (by capacity), 1 (always) or 2 (never)");
    if (c->claim_segments > kClaimSegs || (c->claim_segments & (c->claim_segments - 1u)))
        return r->fail(WF_ERROR_INVALID_ARGUMENT, "claim_segments must be 0 or a power of two up to 64");
    if (c->cost_order > 2) return r->fail(WF_ERROR_INVALID_ARGUMENT, "cost_order must be 0 (default), 1 (on) or 2 (off)");
    const ExtendConfig old = r->ext_cfg;
    r->ext_cfg.block = c->workgroup_size;
    r->ext_cfg.lds_stack = c->lds_stack_entries;
    r->ext_cfg.skip_popped = c->skip_popped_interior ? 1u : 0u;
    r->ext_cfg.stack_limit = c->stack_limit;
    r->ext_cfg.refill_min = c->refill_min ? c->refill_min : ExtendConfig{}.refill_min;
    r->ext_cfg.tlas_lds = c->tlas_in_lds ? 1u : 0u;
    r->ext_cfg.visits = c->interior_visits ? c->interior_visits : ExtendConfig{}.visits;
    r->ext_cfg.layout = c->bvh_layout;
    r->ext_cfg.big = c->big_claims;
    r->ext_cfg.claim_segs = c->claim_segments ? c->claim_segments : kClaimSegs;
    r->ext_cfg.cost_order = c->cost_order == 0 ? ExtendConfig{}.cost_order : (c->cost_order == 1 ? 1u : 0u);
    if (r->ext_cfg.layout != old.layout) r->tlas_valid = false;   // instance records point at the layout's arrays
    r->drop_graph();
    const wf_status s = r->ensure_trace_scratch();
    if (s != WF_OK) { r->ext_cfg = old; return s; }
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_get_traversal_config(const wf_renderer* r, wf_traversal_config* out)
{
    if (!r || !out) return WF_ERROR_INVALID_ARGUMENT;
    out->workgroup_size = r->ext_cfg.block;
    out->lds_stack_entries = r->ext_cfg.lds_stack;
    out->skip_popped_interior = r->ext_cfg.skip_popped;
    out->bvh_layout = r->ext_cfg.layout;
    out->stack_limit = r->ext_cfg.stack_limit;
    out->refill_min = r->ext_cfg.refill_min;
    out->tlas_in_lds = r->ext_cfg.tlas_lds;
    out->interior_visits = r->ext_cfg.visits;
    out->big_claims = r->ext_cfg.big;
    out->claim_segments = r->ext_cfg.claim_segs;
    out->cost_order = r->ext_cfg.cost_order ? 1u : 2u;
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_set_blas_builder(wf_renderer* r, uint32_t builder)
{
    if (!r) return WF_ERROR_INVALID_ARGUMENT;
    if (builder > WF_BLAS_BUILDER_DEVICE) return r->fail(WF_ERROR_INVALID_ARGUMENT, "unknown BLAS builder %u", builder);
    r->blas_builder = builder;
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_set_tlas_builder(wf_renderer* r, uint32_t builder)
{
    if (!r) return WF_ERROR_INVALID_ARGUMENT;
    if (builder > WF_BLAS_BUILDER_DEVICE) return r->fail(WF_ERROR_INVALID_ARGUMENT, "unknown TLAS builder %u", builder);
    r->tlas_builder = builder;
    r->tlas_valid = false;
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_set_tlas_metric(wf_renderer* r, uint32_t metric)
{
    if (!r) return WF_ERROR_INVALID_ARGUMENT;
    if (metric > WF_TLAS_METRIC_UNION) return r->fail(WF_ERROR_INVALID_ARGUMENT, "unknown TLAS metric %u", metric);
    r->tlas_metric = metric;
    r->tlas_valid = false;
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_set_build_options(wf_renderer* r, const wf_build_options* o)
{
    if (!r || !o) return WF_ERROR_INVALID_ARGUMENT;
    if (o->interval_count < 2 || o->interval_count > 1024) return r->fail(WF_ERROR_INVALID_ARGUMENT, "interval_count must be in [2, 1024]");
    r->build_opts = *o;
    return WF_OK;
}

// ---------------------------------------------------------------------------
// resources
// ---------------------------------------------------------------------------

extern "C" WF_API wf_status wf_renderer_create_render_texture(wf_renderer* r, const wf_render_texture_params* p, wf_texture_handle* out)
{
    if (!r || !p || !out || !p->ptr_data || p->width == 0 || p->height == 0) return WF_ERROR_INVALID_ARGUMENT;
    if ((uint64_t)p->width * p->height > (1ull << 28)) return r->fail(WF_ERROR_CAPACITY, "texture too large");
    const size_t texels = (size_t)p->width * p->height;
    std::vector<float> rgba(texels * 4);
    auto srgb_to_linear = [](float c) {   // common.hlsl:261-269
        return c <= 0.04045f ? c / 12.92f : std::pow((c + 0.055f) / 1.055f, 2.4f);
    };
    switch (p->format) {
    case WF_TEXTURE_FORMAT_RGBA8:
    case WF_TEXTURE_FORMAT_RGBA8_SRGB: {
        const uint8_t* s = (const uint8_t*)p->ptr_data;
        for (size_t i = 0; i < texels * 4; ++i) {
            float c = (float)s[i] / 255.0f;
            if (p->format == WF_TEXTURE_FORMAT_RGBA8_SRGB && (i & 3) != 3) c = srgb_to_linear(c);
            rgba[i] = c;
        }
        break;
    }
    case WF_TEXTURE_FORMAT_RG8: {   // two unorm channels; blue 0 and alpha 1 as a sampled RG8 texture reads
        const uint8_t* s = (const uint8_t*)p->ptr_data;
        for (size_t i = 0; i < texels; ++i) {
            rgba[4 * i] = (float)s[2 * i] / 255.0f;
            rgba[4 * i + 1] = (float)s[2 * i + 1] / 255.0f;
            rgba[4 * i + 2] = 0.0f;
            rgba[4 * i + 3] = 1.0f;
        }
        break;
    }
    case WF_TEXTURE_FORMAT_RGB32F: {
        const float* s = (const float*)p->ptr_data;
        for (size_t i = 0; i < texels; ++i) { rgba[4 * i] = s[3 * i]; rgba[4 * i + 1] = s[3 * i + 1]; rgba[4 * i + 2] = s[3 * i + 2]; rgba[4 * i + 3] = 1.0f; }
        break;
    }
    case WF_TEXTURE_FORMAT_RGBA32F:
        std::memcpy(rgba.data(), p->ptr_data, texels * 16);
        break;
    default:
        return r->fail(WF_ERROR_INVALID_ARGUMENT, "unsupported texture format %u", p->format);
    }
    for (float v : rgba) if (!std::isfinite(v)) return r->fail(WF_ERROR_INVALID_ARGUMENT, "texture holds a non-finite value");
    if (r->textures.size() >= kMaxTextures) return r->fail(WF_ERROR_CAPACITY, "too many textures");
    Texture t;
    t.alive = true;
    t.version = 1;
    t.width = p->width;
    t.height = p->height;
    hipError_t e = t.texels.ensure(texels * 16);
    if (e == hipSuccess) e = hipMemcpy(t.texels.p, rgba.data(), texels * 16, hipMemcpyHostToDevice);
    if (e != hipSuccess) { t.texels.release(); return r->hip_fail(e, "texture upload"); }
    r->textures.push_back(std::move(t));
    r->textures_dirty = true;
    *out = make_handle((uint32_t)r->textures.size() - 1, 1);
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_create_render_mesh(wf_renderer* r, const wf_render_mesh_params* p, wf_mesh_handle* out)
{
    if (!r || !p || !out || !p->vertices || !p->indices) return WF_ERROR_INVALID_ARGUMENT;
    if (p->index_count == 0 || p->index_count % 3 != 0) return r->fail(WF_ERROR_INVALID_ARGUMENT, "index_count must be a positive multiple of 3");
    const uint32_t n = p->index_count / 3;
    if (n > kMaxBlasTriangles) return r->fail(WF_ERROR_CAPACITY, "mesh has %u triangles (limit %u)", n, kMaxBlasTriangles);
    for (uint32_t i = 0; i < p->index_count; ++i)
        if (p->indices[i] >= p->vertex_count) return r->fail(WF_ERROR_INVALID_ARGUMENT, "index %u out of range", i);

    // de-index into triangles (renderer.cpp:1251-1258)
    Mesh m;
    m.pos9.resize(9ull * n);
    std::vector<GpuShadeTri> shade(n);
    for (uint32_t t = 0; t < n; ++t) {
        for (int v = 0; v < 3; ++v) {
            const wf_vertex& vx = p->vertices[p->indices[3 * t + v]];
            std::memcpy(&m.pos9[9ull * t + 3 * v], vx.position, 12);
            std::memcpy(shade[t].p[v], vx.position, 12);
            std::memcpy(shade[t].n[v], vx.normal, 12);
            std::memcpy(shade[t].uv[v], vx.uv, 8);
        }
    }
    const uint32_t bins = r->build_opts.interval_count;
    const bool bins_ok = bins >= 2 && bins <= kDevBuildMaxBins;
    if (r->blas_builder == WF_BLAS_BUILDER_DEVICE && !bins_ok)
        return r->fail(WF_ERROR_INVALID_ARGUMENT, "the device BLAS builder takes 2..%u intervals", kDevBuildMaxBins);
    if (r->blas_builder == WF_BLAS_BUILDER_DEVICE || (r->blas_builder == WF_BLAS_BUILDER_AUTO && bins_ok && n >= kAutoDeviceBuildMin)) {
        // build on the GPU: positions up, nodes and leaf-ordered triangles made in place
        hipError_t e = m.nodes.ensure(sizeof(wf_bvh_node) * 2ull * n);
        if (e == hipSuccess) e = m.tris.ensure(sizeof(GpuTri) * n);
        if (e == hipSuccess) e = m.shade.ensure(sizeof(GpuShadeTri) * n);
        if (e == hipSuccess) e = r->build_pos.ensure(36ull * n);
        if (e == hipSuccess) e = r->build_idx.ensure(4ull * n);
        if (e == hipSuccess) e = hipMemcpyAsync(r->build_pos.p, m.pos9.data(), 36ull * n, hipMemcpyHostToDevice, r->stream);
        if (e != hipSuccess) { m.nodes.release(); m.tris.release(); m.shade.release(); return r->hip_fail(e, "mesh upload"); }
        DevBuildResult res;
        std::string berr;
        const wf_status s = device_build_blas((const float*)r->build_pos.p, n, r->build_opts, (wf_bvh_node*)m.nodes.p,
                                              (uint32_t*)r->build_idx.p, r->build_scratch, r->stream, res, berr);
        if (s != WF_OK) { m.nodes.release(); m.tris.release(); m.shade.release(); return r->fail(s, "device BLAS build: %s", berr.c_str()); }
        e = launch_tris_from_index((const float*)r->build_pos.p, (const uint32_t*)r->build_idx.p, n, (GpuTri*)m.tris.p, r->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(m.shade.p, shade.data(), sizeof(GpuShadeTri) * n, hipMemcpyHostToDevice, r->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
        if (e != hipSuccess) { m.nodes.release(); m.tris.release(); m.shade.release(); return r->hip_fail(e, "mesh upload"); }
        m.tri_count = n;
        m.node_count = res.node_count;
        m.max_leaf = res.max_leaf;
        const wf_status es = r->encode_mesh_refs(m);
        if (es != WF_OK) { m.nodes.release(); m.tris.release(); m.shade.release(); return es; }
        std::memcpy(m.blas_min, res.bmin, 12);   // renderer.cpp:107-111
        std::memcpy(m.blas_max, res.bmax, 12);
        m.alive = true;
        m.version = 1;
        r->meshes.push_back(std::move(m));
        *out = make_handle((uint32_t)r->meshes.size() - 1, 1);
        return WF_OK;
    }
    BlasData b;
    wf_status s = build_blas(m.pos9.data(), n, r->build_opts, b);
    if (s != WF_OK) return r->fail(s, "BLAS build failed (non-finite vertex or bad options)");

    // device layouts: nodes as built, triangles permuted into leaf order
    std::vector<GpuTri> perm(n);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t ti = b.idx[i];
        const float* q = &m.pos9[9ull * ti];
        GpuTri& g = perm[i];
        for (int k = 0; k < 3; ++k) {
            g.v0[k] = q[k];
            g.e1[k] = q[3 + k] - q[k];
            g.e2[k] = q[6 + k] - q[k];
        }
        g.prim = ti;
        g.pad0 = g.pad1 = 0;
    }
    hipError_t e = m.nodes.ensure(sizeof(wf_bvh_node) * b.node_count);
    if (e == hipSuccess) e = m.tris.ensure(sizeof(GpuTri) * n);
    if (e == hipSuccess) e = m.shade.ensure(sizeof(GpuShadeTri) * n);
    if (e == hipSuccess) e = hipMemcpy(m.nodes.p, b.nodes.data(), sizeof(wf_bvh_node) * b.node_count, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m.tris.p, perm.data(), sizeof(GpuTri) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m.shade.p, shade.data(), sizeof(GpuShadeTri) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) { m.nodes.release(); m.tris.release(); m.shade.release(); return r->hip_fail(e, "mesh upload"); }
    m.tri_count = n;
    m.node_count = b.node_count;
    m.max_leaf = b.max_leaf;
    s = r->encode_mesh_refs(m);
    if (s != WF_OK) { m.nodes.release(); m.tris.release(); m.shade.release(); return s; }
    std::memcpy(m.blas_min, b.nodes[0].aabb_min, 12);   // renderer.cpp:107-111
    std::memcpy(m.blas_max, b.nodes[0].aabb_max, 12);
    m.alive = true;
    m.version = 1;
    r->meshes.push_back(std::move(m));
    *out = make_handle((uint32_t)r->meshes.size() - 1, 1);
    return WF_OK;
}

// The renderer's node copy in the traversal's form: every leaf's left_first
// word becomes its reference (kernel k_encode_blas_refs), checked in range.
wf_status wf_renderer::encode_mesh_refs(wf::Mesh& m)
{
    hipError_t e = d_check.ensure(4);
    if (e == hipSuccess) e = launch_fill_u32((uint32_t*)d_check.p, 0u, 1, stream);
    if (e == hipSuccess)
        e = launch_encode_blas_refs((wf_bvh_node*)m.nodes.p, m.node_count, (GpuTri*)m.tris.p, m.tri_count, (uint32_t*)d_check.p, stream);
    uint32_t bad = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&bad, d_check.p, 4, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&m.root_ref, (const char*)m.nodes.p + 12, 4, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "BLAS reference encoding");
    if (bad) return fail(WF_ERROR_INVALID_STATE, "a BLAS leaf's triangle range is out of bounds");
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_submit_render_mesh(wf_renderer* r, wf_mesh_handle mh, const float transform[16], const wf_material_asset* mat)
{
    if (!r || !transform || !mat) return WF_ERROR_INVALID_ARGUMENT;
    Mesh* m = r->find_mesh(mh);
    if (!m) return r->fail(WF_ERROR_NOT_FOUND, "mesh handle {index %u, version %u} is not valid", handle_index(mh), handle_version(mh));
    if (!r->scene_open || r->cur_slot >= r->slots.size())
        return r->fail(WF_ERROR_INVALID_STATE, "submit_render_mesh outside begin_scene / end_scene");
    size_t total = 0;
    for (const auto& sl : r->slots) total += sl.instances.size();
    if (total >= kMaxInstances) return r->fail(WF_ERROR_CAPACITY, "exceeded capacity of instances (%u)", kMaxInstances);
    for (int i = 0; i < 16; ++i) if (!std::isfinite(transform[i])) return r->fail(WF_ERROR_INVALID_ARGUMENT, "non-finite transform");
    InstanceRec rec;
    rec.mesh = handle_index(mh);
    std::memcpy(rec.transform, transform, 64);
    rec.material = *mat;
    // software instance record (renderer.cpp:1318-1336)
    make_bvh_instance(transform, m->blas_min, m->blas_max, rec.mesh, rec.sw);
    r->slots[r->cur_slot].instances.push_back(rec);
    r->tlas_valid = false;
    return WF_OK;
}

// ---------------------------------------------------------------------------
// TLAS build + upload (renderer.cpp:637-670), one TLAS per scene slot
// ---------------------------------------------------------------------------

// Every slot's TLAS is built by the reference algorithm, then the slots are
// concatenated: child indices and instance indices are offset by the slot's
// node and instance bases (both halves of left_right stay 16-bit), so slot 0
// is unchanged and slot s is entered at its root node tlas_root.
wf_status wf_renderer::upload_scene()
{
    if (tlas_valid) return WF_OK;
    if (slots.empty()) return fail(WF_ERROR_INVALID_STATE, "no scene (call begin_scene)");
    if (ext_cfg.layout == 1) {
        for (auto& sl : slots)
            for (const InstanceRec& rec : sl.instances) {
                Mesh& m = meshes[rec.mesh];
                if (m.quads.p) continue;
                hipError_t e = m.quads.ensure((size_t)kQuadBytes * (m.node_count / 2u));
                if (e == hipSuccess) e = launch_make_quads((const wf_bvh_node*)m.nodes.p, m.node_count, m.quads.p, stream);
                if (e != hipSuccess) return hip_fail(e, "4-wide BLAS records");
            }
    }
    std::vector<wf_tlas_node> all_nodes;
    std::vector<GpuInstance> gi;
    std::vector<GpuShadeInstance> si;
    for (auto& sl : slots) {
        if (sl.instances.empty()) return fail(WF_ERROR_INVALID_STATE, "a scene has no instances submitted");
        if (sl.instances.size() > kMaxTlasInstances)
            return fail(WF_ERROR_CAPACITY, "%zu instances exceed the TLAS limit %u", sl.instances.size(), kMaxTlasInstances);
        std::vector<wf_bvh_instance> sw(sl.instances.size());
        for (size_t i = 0; i < sl.instances.size(); ++i) sw[i] = sl.instances[i].sw;
        TlasData t;
        const uint32_t m = (uint32_t)sw.size();
        if (tlas_builder == WF_BLAS_BUILDER_DEVICE || (tlas_builder == WF_BLAS_BUILDER_AUTO && m >= kAutoDeviceTlasMin)) {
            hipError_t e = tbuild_inst.ensure(sizeof(wf_bvh_instance) * m);
            if (e == hipSuccess) e = tbuild_nodes.ensure(sizeof(wf_tlas_node) * 2ull * m);
            if (e == hipSuccess) e = hipMemcpyAsync(tbuild_inst.p, sw.data(), sizeof(wf_bvh_instance) * m, hipMemcpyHostToDevice, stream);
            if (e != hipSuccess) return hip_fail(e, "TLAS instance upload");
            uint32_t count = 0;
            std::string berr;
            const wf_status s = device_build_tlas((const wf_bvh_instance*)tbuild_inst.p, m, (wf_tlas_node*)tbuild_nodes.p, tbuild_list,
                                                  tbuild_ids, tbuild_status, stream, count, berr, tlas_metric);
            if (s != WF_OK) return fail(s, "device TLAS build: %s", berr.c_str());
            t.nodes.resize(count);
            e = hipMemcpyAsync(t.nodes.data(), tbuild_nodes.p, sizeof(wf_tlas_node) * count, hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            if (e != hipSuccess) return hip_fail(e, "TLAS readback");
            t.node_count = count;
            t.instances = sw;
        } else {
            wf_status s = build_tlas(sw.data(), m, t, tlas_metric);
            if (s != WF_OK) return fail(s, "TLAS build failed");
        }
        const uint32_t node_base = (uint32_t)all_nodes.size();
        const uint32_t inst_base = (uint32_t)gi.size();
        if (node_base + t.node_count > 0xFFFFu)
            return fail(WF_ERROR_CAPACITY, "the scenes' TLAS nodes exceed 65535 (16-bit child indices)");
        sl.tlas_root = node_base;
        sl.inst_base = inst_base;
        std::memcpy(sl.bounds_min, t.nodes[0].aabb_min, 12);
        std::memcpy(sl.bounds_max, t.nodes[0].aabb_max, 12);
        for (uint32_t k = 0; k < t.node_count; ++k) {
            wf_tlas_node nd = t.nodes[k];
            if (nd.left_right == 0) nd.instance_idx += inst_base;
            else nd.left_right = ((nd.left_right & 0xFFFFu) + node_base) + (((nd.left_right >> 16) + node_base) << 16);
            all_nodes.push_back(nd);
        }
        for (const InstanceRec& rec : sl.instances) {
            const Mesh& m = meshes[rec.mesh];
            GpuInstance g;
            for (int c = 0; c < 4; ++c)
                for (int rr = 0; rr < 3; ++rr) g.m[c * 3 + rr] = rec.sw.world_to_local[c * 4 + rr];
            g.nodes = (const wf_bvh_node*)m.nodes.p;
            g.tris = (const GpuTri*)m.tris.p;
            g.quads = (const char*)m.quads.p;
            // 4-wide record 0 holds the root alone, an interior reference
            // there being a record number (left_first / 2)
            g.root2 = m.root_ref;
            g.root4 = (m.root_ref & kLeafBit) ? m.root_ref : m.root_ref >> 1;
            gi.push_back(g);
            GpuShadeInstance sh;
            std::memset(&sh, 0, sizeof(sh));
            std::memcpy(sh.local_to_world, rec.transform, 64);
            for (int k = 0; k < 3; ++k) {
                sh.base_color[k] = rec.material.base_color_factor[k];
                sh.emissive_factor[k] = rec.material.emissive_factor[k];
            }
            sh.emissive_strength = rec.material.emissive_strength;
            const Texture* bt = find_texture(rec.material.base_color_texture);
            const Texture* et = find_texture(rec.material.emissive_texture);
            sh.base_color_tex = bt ? (int32_t)handle_index(rec.material.base_color_texture) : -1;
            sh.emissive_tex = et ? (int32_t)handle_index(rec.material.emissive_texture) : -1;
            sh.normal_tex = find_texture(rec.material.normal_texture) ? (int32_t)handle_index(rec.material.normal_texture) : -1;
            sh.metal_rough_tex = find_texture(rec.material.metallic_roughness_texture)
                                     ? (int32_t)handle_index(rec.material.metallic_roughness_texture) : -1;
            sh.tris = (const GpuShadeTri*)m.shade.p;
            si.push_back(sh);
        }
    }
    // child-pair records of the interior TLAS nodes, numbered densely (see
    // TraceScene); the device node copy carries each interior node's number
    std::vector<uint32_t> rec_of(all_nodes.size(), 0);
    uint32_t n_rec = 0;
    for (size_t k = 0; k < all_nodes.size(); ++k)
        if (all_nodes[k].left_right != 0) rec_of[k] = n_rec++;
    std::vector<wf_tlas_node> dev_nodes = all_nodes;
    for (size_t k = 0; k < all_nodes.size(); ++k)
        if (all_nodes[k].left_right != 0) dev_nodes[k].instance_idx = rec_of[k];
    std::vector<float> pairs(16 * std::max<uint32_t>(n_rec, 1), 0.0f);
    for (size_t k = 0; k < all_nodes.size(); ++k) {
        const uint32_t lr = all_nodes[k].left_right;
        if (lr == 0) continue;
        const uint32_t child[2] = {lr >> 16, lr & 0xFFFFu};
        for (int c = 0; c < 2; ++c) {
            const wf_tlas_node& cn = all_nodes[child[c]];
            float* slot = &pairs[16 * rec_of[k] + 8 * c];
            const bool leaf = cn.left_right == 0;
            const uint32_t ref = leaf ? encode_ref(cn.instance_idx, 1u) : 2u * rec_of[child[c]];
            const uint32_t kind = leaf ? 1u : 0u;
            std::memcpy(slot, cn.aabb_min, 12);
            std::memcpy(slot + 3, &ref, 4);
            std::memcpy(slot + 4, cn.aabb_max, 12);
            std::memcpy(slot + 7, &kind, 4);
        }
    }
    // 4-wide TLAS records: record k = pair k, then the pairs of its interior
    // children; interior references become record numbers (see kQuadBytes)
    std::vector<float> quads;
    if (ext_cfg.layout == 1) {
        quads.assign(48 * std::max<uint32_t>(n_rec, 1), 0.0f);
        auto put = [](float* dst, const float* src) {
            std::memcpy(dst, src, 32);
            uint32_t ref, kind;
            std::memcpy(&ref, src + 3, 4);
            std::memcpy(&kind, src + 7, 4);
            if (kind == 0) { ref >>= 1; std::memcpy(dst + 3, &ref, 4); }
        };
        for (uint32_t k = 0; k < n_rec; ++k) {
            float* q = &quads[48 * (size_t)k];
            for (int c = 0; c < 2; ++c) {
                const float* src = &pairs[16 * (size_t)k + 8 * c];
                put(q + 8 * c, src);
                uint32_t ref, kind;
                std::memcpy(&ref, src + 3, 4);
                std::memcpy(&kind, src + 7, 4);
                if (kind == 0) {
                    put(q + 16 + 16 * c, &pairs[16 * (size_t)(ref >> 1)]);
                    put(q + 24 + 16 * c, &pairs[16 * (size_t)(ref >> 1) + 8]);
                }
            }
        }
    }
    drop_graph();
    hipError_t e = d_tnodes.ensure(sizeof(wf_tlas_node) * all_nodes.size());
    if (e == hipSuccess && !quads.empty()) e = d_tquads.ensure(sizeof(float) * quads.size());
    if (e == hipSuccess && !quads.empty())
        e = hipMemcpyAsync(d_tquads.p, quads.data(), sizeof(float) * quads.size(), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = d_tpairs.ensure(sizeof(float) * pairs.size());
    if (e == hipSuccess) e = hipMemcpyAsync(d_tpairs.p, pairs.data(), sizeof(float) * pairs.size(), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = d_inst.ensure(sizeof(GpuInstance) * gi.size());
    if (e == hipSuccess) e = d_shade_inst.ensure(sizeof(GpuShadeInstance) * si.size());
    if (e == hipSuccess) e = hipMemcpyAsync(d_tnodes.p, dev_nodes.data(), sizeof(wf_tlas_node) * dev_nodes.size(), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_inst.p, gi.data(), sizeof(GpuInstance) * gi.size(), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_shade_inst.p, si.data(), sizeof(GpuShadeInstance) * si.size(), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "TLAS upload");
    n_tpairs = n_rec;
    if (ext_grid) {
        const wf_status ts = update_tlds_grid();
        if (ts != WF_OK) return ts;
    }
    wf_status s = upload_materials_and_emitters();
    if (s != WF_OK) return s;
    tlas_valid = true;
    return WF_OK;
}

extern "C" WF_API wf_status wf_renderer_build_tlas(wf_renderer* r)
{
    if (!r) return WF_ERROR_INVALID_ARGUMENT;
    r->tlas_valid = false;
    return r->upload_scene();
}

extern "C" WF_API wf_status wf_renderer_get_scene_bounds(const wf_renderer* r, float out_min[3], float out_max[3])
{
    if (!r || !out_min || !out_max) return WF_ERROR_INVALID_ARGUMENT;
    if (!r->tlas_valid || r->slots.empty()) return WF_ERROR_INVALID_STATE;
    std::memcpy(out_min, r->slots[0].bounds_min, 12);
    std::memcpy(out_max, r->slots[0].bounds_max, 12);
    return WF_OK;
}

// ---------------------------------------------------------------------------
// extend on caller rays
// ---------------------------------------------------------------------------

wf_status wf_renderer::run_extend(const wf_ray* d_rays, const uint32_t* d_count, uint32_t n, wf_hit* d_hits,
                                  uint32_t* d_visible, uint32_t* d_counters, bool any, const uint32_t* d_roots,
                                  const OcclusionQueue* second, bool reset_overflow, const uint32_t* d_order)
{
    if (d_order && (second || d_roots)) return fail(WF_ERROR_INVALID_ARGUMENT, "a cost-ordered queue has one scene and one queue");
    if (second && (any || d_counters)) return fail(WF_ERROR_INVALID_ARGUMENT, "a second queue needs a closest-hit launch");
    // the queue cursors start at zero; word 1 counts overflows, cleared here
    // only when the caller asks (a frame counts them over all of its launches)
    hipError_t e = launch_clear_claims((uint32_t*)d_work.p, reset_overflow, stream);
    if (e != hipSuccess) return hip_fail(e, "work counter reset");
    ExtendLaunch L{};
    L.scene.tnodes = (const wf_tlas_node*)d_tnodes.p;
    L.scene.tpairs = (const float4*)d_tpairs.p;
    L.scene.n_tpairs = n_tpairs;
    L.scene.inst = (const GpuInstance*)d_inst.p;
    L.scene.tquads = (const float4*)d_tquads.p;
    L.rays = d_rays;
    L.order = d_order;
    L.ray_root = d_roots;
    L.d_count = d_count;
    L.n = n;
    L.hits = d_hits;
    L.visible = d_visible;
    L.counters = d_counters;
    L.work = (uint32_t*)d_work.p;
    L.spill = (uint2*)d_spill.p;
    L.spill_per_thread = spill_per_thread;
    L.stack_cap = ext_cfg.stack_limit ? std::min(ext_cfg.stack_limit, kStackTotal) : kStackTotal;
    L.grid = ext_grid;
    L.grid_tlds = ext_grid_tlds;
    L.any_hit = any ? 1 : 0;
    L.refill_min = ext_cfg.refill_min;
    L.visits = ext_cfg.visits;
    L.claim_segs = ext_cfg.claim_segs;
    if (second) {
        L.rays2 = second->rays;
        L.ray_root2 = second->roots;
        L.d_count2 = second->d_count;
        L.n2 = 0;
        L.visible2 = second->visible;
    }
    e = launch_extend(L, ext_cfg, stream);
    if (e != hipSuccess) return hip_fail(e, "extend launch");
    return WF_OK;
}

wf_status wf_renderer::cost_order(const wf_ray* d_rays, uint32_t n, const uint32_t** out)
{
    *out = nullptr;
    // a queue no longer than two rays per persistent lane starts (nearly)
    // every ray at once: there is no order to choose
    if (!ext_cfg.cost_order || slots.empty() || (uint64_t)n < 2ull * ext_grid * ext_cfg.block) return WF_OK;
    hipError_t e = d_order.ensure(4ull * n);
    if (e == hipSuccess) e = d_order_scratch.ensure(4ull * order_scratch_words(n));
    if (e != hipSuccess) return hip_fail(e, "cost-ordered queue allocation");
    e = launch_cost_order(d_rays, n, slots[0].bounds_min, slots[0].bounds_max, (uint32_t*)d_order_scratch.p,
                          (uint32_t*)d_order.p, stream);
    if (e != hipSuccess) return hip_fail(e, "cost-ordered queue");
    *out = (const uint32_t*)d_order.p;
    return WF_OK;
}

wf_status wf_renderer::reset_overflows()
{
    hipError_t e = launch_fill_u32((uint32_t*)d_work.p, 0u, 16, stream);
    if (e != hipSuccess) return hip_fail(e, "overflow counter reset");
    return WF_OK;
}

wf_status wf_renderer::read_overflows(uint32_t* out)
{
    hipError_t e = hipMemcpyAsync(out, (uint32_t*)d_work.p + 1, 4, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "overflow counter readback");
    return WF_OK;
}

static wf_status trace_host(wf_renderer* r, const wf_ray* rays, uint32_t n, wf_hit* hits, uint32_t* visible, uint32_t* counters)
{
    if (!r || (n && !rays) || (n && !hits && !visible)) return WF_ERROR_INVALID_ARGUMENT;
    wf_status s = r->upload_scene();
    if (s != WF_OK) return s;
    if (n == 0) return WF_OK;
    const bool any = visible != nullptr;
    const size_t out_bytes = any ? 4ull * n : sizeof(wf_hit) * (size_t)n;
    hipError_t e = r->d_tmp_rays.ensure(sizeof(wf_ray) * (size_t)n);
    if (e == hipSuccess) e = r->d_tmp_hits.ensure(out_bytes);
    if (e == hipSuccess && counters) e = r->d_tmp_aux.ensure(16ull * n);
    if (e == hipSuccess) e = hipMemcpyAsync(r->d_tmp_rays.p, rays, sizeof(wf_ray) * (size_t)n, hipMemcpyHostToDevice, r->stream);
    if (e != hipSuccess) return r->hip_fail(e, "ray upload");
    if ((s = r->reset_overflows()) != WF_OK) return s;
    const uint32_t* order = nullptr;
    if ((s = r->cost_order((const wf_ray*)r->d_tmp_rays.p, n, &order)) != WF_OK) return s;
    s = r->run_extend((const wf_ray*)r->d_tmp_rays.p, nullptr, n, any ? nullptr : (wf_hit*)r->d_tmp_hits.p,
                      any ? (uint32_t*)r->d_tmp_hits.p : nullptr, counters ? (uint32_t*)r->d_tmp_aux.p : nullptr, any,
                      nullptr, nullptr, false, order);
    if (s != WF_OK) return s;
    e = hipMemcpyAsync(any ? (void*)visible : (void*)hits, r->d_tmp_hits.p, out_bytes, hipMemcpyDeviceToHost, r->stream);
    if (e == hipSuccess && counters) e = hipMemcpyAsync(counters, r->d_tmp_aux.p, 16ull * n, hipMemcpyDeviceToHost, r->stream);
    if (e != hipSuccess) return r->hip_fail(e, "hit readback");
    uint32_t ov = 0;
    s = r->read_overflows(&ov);
    if (s != WF_OK) return s;
    if (ov) return r->fail(WF_ERROR_STACK_OVERFLOW, "%u rays overflowed the traversal stack", ov);
    return WF_OK;
}

extern "C" WF_API wf_status wf_trace_closest(wf_renderer* r, const wf_ray* rays, uint32_t n, wf_hit* hits)
{
    if (!hits && n) return WF_ERROR_INVALID_ARGUMENT;
    return trace_host(r, rays, n, hits, nullptr, nullptr);
}

extern "C" WF_API wf_status wf_trace_occluded(wf_renderer* r, const wf_ray* rays, uint32_t n, uint32_t* visible)
{
    if (!visible && n) return WF_ERROR_INVALID_ARGUMENT;
    return trace_host(r, rays, n, nullptr, visible, nullptr);
}

extern "C" WF_API wf_status wf_trace_closest_counters(wf_renderer* r, const wf_ray* rays, uint32_t n, wf_hit* hits, uint32_t* counters)
{
    if ((!hits || !counters) && n) return WF_ERROR_INVALID_ARGUMENT;
    return trace_host(r, rays, n, hits, nullptr, counters);
}

extern "C" WF_API wf_status wf_trace_closest_device(wf_renderer* r, const wf_ray* d_rays, uint32_t n, wf_hit* d_hits)
{
    if (!r || (n && (!d_rays || !d_hits))) return WF_ERROR_INVALID_ARGUMENT;
    wf_status s = r->upload_scene();
    if (s != WF_OK) return s;
    if (n == 0) return r->reset_overflows();
    const uint32_t* order = nullptr;
    if ((s = r->cost_order(d_rays, n, &order)) != WF_OK) return s;
    return r->run_extend(d_rays, nullptr, n, d_hits, nullptr, nullptr, false, nullptr, nullptr, true, order);
}

extern "C" WF_API wf_status wf_renderer_read_stack_overflows(wf_renderer* r, uint32_t* out)
{
    if (!r || !out) return WF_ERROR_INVALID_ARGUMENT;
    if (!r->d_work.p) { *out = 0; return WF_OK; }
    return r->read_overflows(out);
}
