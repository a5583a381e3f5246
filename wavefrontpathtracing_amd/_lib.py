"""ctypes binding of libwf.so (include/wf.h, include/wf_scene.h).

The native library is required: importing this module on a machine where it
is not built raises immediately.  There is no Python or CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# WF_LIB selects another build of the library (the sanitised host build of
# build.build_sanitized, for instance); by default the in-tree libwf.so
LIB_PATH = os.environ.get("WF_LIB") or os.path.join(_HERE, "libwf.so")

# ---- wire formats (numpy) ---------------------------------------------------
VERTEX = np.dtype([("position", "<f4", 3), ("normal", "<f4", 3), ("tangent", "<f4", 4), ("uv", "<f4", 2)])
RAY = np.dtype([("origin", "<f4", 3), ("tmin", "<f4"), ("direction", "<f4", 3), ("tmax", "<f4")])
HIT = np.dtype([("instance_idx", "<u4"), ("primitive_idx", "<u4"), ("t", "<f4"), ("bary", "<f4", 2)])
BVH_NODE = np.dtype([("aabb_min", "<f4", 3), ("left_first", "<u4"), ("aabb_max", "<f4", 3), ("prim_count", "<u4")])
BVH_TRIANGLE = np.dtype([("p0", "<f4", 3), ("p1", "<f4", 3), ("p2", "<f4", 3)])
TLAS_NODE = np.dtype([("aabb_min", "<f4", 3), ("left_right", "<u4"), ("aabb_max", "<f4", 3), ("instance_idx", "<u4")])
BVH_INSTANCE = np.dtype([("world_to_local", "<f4", 16), ("aabb_min", "<f4", 3), ("aabb_max", "<f4", 3), ("bvh_index", "<u4")])
assert VERTEX.itemsize == 48 and RAY.itemsize == 32 and HIT.itemsize == 20
assert BVH_NODE.itemsize == 32 and TLAS_NODE.itemsize == 32 and BVH_INSTANCE.itemsize == 92

INVALID_HANDLE = (1 << 64) - 1
MISS = 0xFFFFFFFF

STATUS = {
    0: "WF_OK", 1: "WF_ERROR_INVALID_ARGUMENT", 2: "WF_ERROR_CAPACITY", 3: "WF_ERROR_OUT_OF_MEMORY",
    4: "WF_ERROR_DEVICE", 5: "WF_ERROR_STACK_OVERFLOW", 6: "WF_ERROR_INVALID_STATE", 7: "WF_ERROR_NOT_FOUND",
}


class WfError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


# ---- ctypes structs ------------------------------------------------------------
class InitParams(C.Structure):
    _fields_ = [("render_width", C.c_uint32), ("render_height", C.c_uint32), ("backbuffer_count", C.c_uint32), ("vsync", C.c_uint32)]


class Camera(C.Structure):
    _fields_ = [("inv_view_matrix", C.c_float * 16), ("view_matrix", C.c_float * 16), ("vfov_deg", C.c_float),
                ("yaw", C.c_float), ("pitch", C.c_float)]


class View(C.Structure):
    _fields_ = [("world_to_view", C.c_float * 16), ("view_to_world", C.c_float * 16), ("view_to_clip", C.c_float * 16),
                ("clip_to_view", C.c_float * 16), ("render_dim", C.c_float * 2), ("near_plane", C.c_float), ("far_plane", C.c_float)]


class RenderTextureParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("bits_per_pixel", C.c_uint32), ("format", C.c_uint32),
                ("ptr_data", C.c_void_p)]


class RenderMeshParams(C.Structure):
    _fields_ = [("vertex_count", C.c_uint32), ("vertices", C.c_void_p), ("index_count", C.c_uint32), ("indices", C.c_void_p)]


class MaterialAsset(C.Structure):
    _fields_ = [("base_color_factor", C.c_float * 4), ("metallic_factor", C.c_float), ("roughness_factor", C.c_float),
                ("emissive_factor", C.c_float * 3), ("emissive_strength", C.c_float),
                ("base_color_texture", C.c_uint64), ("normal_texture", C.c_uint64),
                ("metallic_roughness_texture", C.c_uint64), ("emissive_texture", C.c_uint64)]


class RenderSettings(C.Structure):
    _fields_ = [("use_wavefront_pathtracing", C.c_uint32), ("use_software_rt", C.c_uint32), ("render_view_mode", C.c_uint32),
                ("max_bounces", C.c_uint32), ("cosine_weighted_diffuse", C.c_uint32), ("accumulate", C.c_uint32),
                ("hdr_env_strength", C.c_float), ("spp", C.c_uint32), ("rr_start_bounce", C.c_uint32),
                ("rr_survival", C.c_float), ("next_event_estimation", C.c_uint32), ("frame_seed", C.c_uint32),
                ("use_hip_graph", C.c_uint32), ("separate_connect", C.c_uint32)]


class BuildOptions(C.Structure):
    _fields_ = [("interval_count", C.c_uint32), ("subdivide_single_prim", C.c_uint32)]


class FrameStats(C.Structure):
    _fields_ = [("ms_clear", C.c_float), ("ms_generate", C.c_float), ("ms_extend", C.c_float), ("ms_shade", C.c_float),
                ("ms_connect", C.c_float), ("ms_post", C.c_float), ("ms_total", C.c_float),
                ("rays_extended", C.c_uint32), ("shadow_rays", C.c_uint32), ("stack_overflows", C.c_uint32),
                ("rays_per_bounce", C.c_uint32 * 9), ("shadow_rays_per_bounce", C.c_uint32 * 9)]


class TraversalConfig(C.Structure):
    _fields_ = [("workgroup_size", C.c_uint32), ("lds_stack_entries", C.c_uint32), ("skip_popped_interior", C.c_uint32),
                ("bvh_layout", C.c_uint32), ("stack_limit", C.c_uint32), ("refill_min", C.c_uint32),
                ("tlas_in_lds", C.c_uint32), ("interior_visits", C.c_uint32), ("big_claims", C.c_uint32),
                ("claim_segments", C.c_uint32), ("cost_order", C.c_uint32)]


class Pcg32(C.Structure):
    _fields_ = [("state", C.c_uint64), ("inc", C.c_uint64)]


# ---- function table ------------------------------------------------------------
_vp = C.c_void_p
_u32 = C.c_uint32
_u64 = C.c_uint64
_f32 = C.c_float
_st = C.c_int

SIGNATURES = {
    "wf_renderer_init": (_st, [C.POINTER(InitParams), C.POINTER(_vp)]),
    "wf_renderer_exit": (_st, [_vp]),
    "wf_renderer_last_error": (C.c_char_p, [_vp]),
    "wf_renderer_get_stream": (_st, [_vp, C.POINTER(_vp)]),
    "wf_renderer_begin_frame": (_st, [_vp]),
    "wf_renderer_end_frame": (_st, [_vp]),
    "wf_renderer_begin_scene": (_st, [_vp, C.POINTER(Camera), _u64]),
    "wf_renderer_render": (_st, [_vp]),
    "wf_renderer_end_scene": (_st, [_vp]),
    "wf_renderer_begin_batch": (_st, [_vp]),
    "wf_renderer_batch_add_scene": (_st, [_vp, C.POINTER(Camera), _u32, _u32, C.POINTER(_u32)]),
    "wf_renderer_get_scene_count": (_st, [_vp, C.POINTER(_u32)]),
    "wf_renderer_render_scenes": (_st, [_vp, _u32, _u32]),
    "wf_renderer_read_scene_energy": (_st, [_vp, _u32, _vp]),
    "wf_renderer_render_adjoint": (_st, [_vp, _vp, _u64, _vp, _vp]),
    "wf_renderer_render_adjoint_device": (_st, [_vp, _vp, _u64, _vp, _vp]),
    "wf_renderer_update_texture": (_st, [_vp, _u64, _vp]),
    "wf_renderer_update_texture_device": (_st, [_vp, _u64, _vp]),
    "wf_renderer_set_emission_strength": (_st, [_vp, _u32, _f32]),
    "wf_renderer_read_energy_device": (_st, [_vp, _vp]),
    "wf_renderer_create_render_texture": (_st, [_vp, C.POINTER(RenderTextureParams), C.POINTER(_u64)]),
    "wf_renderer_create_render_mesh": (_st, [_vp, C.POINTER(RenderMeshParams), C.POINTER(_u64)]),
    "wf_renderer_submit_render_mesh": (_st, [_vp, _u64, _vp, C.POINTER(MaterialAsset)]),
    "wf_renderer_get_settings": (_st, [_vp, C.POINTER(RenderSettings)]),
    "wf_renderer_set_settings": (_st, [_vp, C.POINTER(RenderSettings)]),
    "wf_renderer_set_build_options": (_st, [_vp, C.POINTER(BuildOptions)]),
    "wf_renderer_set_traversal_config": (_st, [_vp, C.POINTER(TraversalConfig)]),
    "wf_renderer_get_traversal_config": (_st, [_vp, C.POINTER(TraversalConfig)]),
    "wf_camera_create": (_st, [C.POINTER(Camera), _vp, _vp, _f32]),
    "wf_renderer_get_view": (_st, [_vp, C.POINTER(View)]),
    "wf_make_view": (_st, [C.POINTER(Camera), _u32, _u32, C.POINTER(View)]),
    "wf_renderer_read_energy": (_st, [_vp, _vp]),
    "wf_renderer_read_accumulator": (_st, [_vp, _vp]),
    "wf_renderer_read_final_color": (_st, [_vp, _vp]),
    "wf_renderer_get_frame_stats": (_st, [_vp, C.POINTER(FrameStats)]),
    "wf_write_pfm": (_st, [C.c_char_p, _vp, _u32, _u32]),
    "wf_write_ppm": (_st, [C.c_char_p, _vp, _u32, _u32]),
    "wf_renderer_build_tlas": (_st, [_vp]),
    "wf_renderer_get_scene_bounds": (_st, [_vp, _vp, _vp]),
    "wf_trace_closest": (_st, [_vp, _vp, _u32, _vp]),
    "wf_trace_closest_device": (_st, [_vp, _vp, _u32, _vp]),
    "wf_renderer_read_stack_overflows": (_st, [_vp, C.POINTER(_u32)]),
    "wf_trace_occluded": (_st, [_vp, _vp, _u32, _vp]),
    "wf_trace_closest_counters": (_st, [_vp, _vp, _u32, _vp, _vp]),
    "wf_bvh_build": (_st, [_vp, _u32, C.POINTER(BuildOptions), _vp, _u64, C.POINTER(_u64), C.POINTER(_u32)]),
    "wf_renderer_set_blas_builder": (_st, [_vp, _u32]),
    "wf_bvh_builder_create": (_st, [C.POINTER(_vp)]),
    "wf_bvh_builder_destroy": (_st, [_vp]),
    "wf_bvh_builder_build_device": (_st, [_vp, _vp, _u32, C.POINTER(BuildOptions), _vp, _u64, _vp, C.POINTER(_u32), _vp]),
    "wf_bvh_builder_last_build": (_st, [_vp, C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32)]),
    "wf_bvh_builder_last_error": (C.c_char_p, [_vp]),
    "wf_tlas_build_device": (_st, [_vp, _u32, _vp, _u64, C.POINTER(_u32), _vp]),
    "wf_tlas_build_device_metric": (_st, [_vp, _u32, _u32, _vp, _u64, C.POINTER(_u32), _vp]),
    "wf_renderer_set_tlas_builder": (_st, [_vp, _u32]),
    "wf_renderer_set_tlas_metric": (_st, [_vp, _u32]),
    "wf_bvh_build_device": (_st, [_vp, _u32, C.POINTER(BuildOptions), _vp, _u64, _vp, C.POINTER(_u32), _vp]),
    "wf_bvh_build_device_noalloc": (_st, [_vp, _u32, C.POINTER(BuildOptions), _vp, _u64, _vp, C.POINTER(_u32), _vp]),
    "wf_selftest_reciprocal": (_st, [C.POINTER(_u64), C.POINTER(_u32)]),
    "wf_selftest_cost_order": (_st, [_vp, _u32, _vp, _vp, _vp, _vp]),
    "wf_tlas_build": (_st, [_vp, _u32, _vp, _u64, C.POINTER(_u64), C.POINTER(_u32)]),
    "wf_tlas_build_metric": (_st, [_vp, _u32, _u32, _vp, _u64, C.POINTER(_u64), C.POINTER(_u32)]),
    "wf_make_bvh_instance": (_st, [_vp, _vp, _vp, _u32, _vp]),
    # wf_scene.h
    "wfs_pcg32_seed": (None, [C.POINTER(Pcg32), _u64, _u64]),
    "wfs_pcg32_next": (_u32, [C.POINTER(Pcg32)]),
    "wfs_pcg32_float": (_f32, [C.POINTER(Pcg32)]),
    "wfs_random_soup": (_st, [_u64, _u32, _f32, _f32, _f32, _vp, _vp]),
    "wfs_heightfield": (_st, [_u64, _u32, _u32, _f32, _vp, _vp]),
    "wfs_icosphere": (_st, [_u32, _vp, _vp, C.POINTER(_u32), C.POINTER(_u32)]),
    "wfs_box_room": (_st, [_vp, _vp]),
    "wfs_quad": (_st, [_vp, _vp]),
    "wfs_random_transform": (None, [C.POINTER(Pcg32), _f32, _f32, _f32, _vp]),
    "wfs_sphere_rays": (_st, [_u64, _vp, _vp, _u32, _vp]),
    "wfs_primary_rays": (_st, [C.POINTER(View), _u32, _u32, _vp]),
    "wfs_value_noise_texture": (_st, [_u64, _u32, _u32, _u32, _f32, _f32, _vp]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libwf.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python __graft_entry__.py` or "
                               "`python -m wavefrontpathtracing_amd.build` first")
        # PyTorch-ROCm ships its own HIP runtime.  Loading torch first makes
        # libwf.so bind to that same runtime (one runtime, one device context
        # per process), so torch tensors, streams and events interoperate with
        # the library; a second runtime loaded behind ours would find no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def check(status: int, renderer=None, what: str = "") -> None:
    if status != 0:
        msg = what
        if renderer is not None:
            m = lib().wf_renderer_last_error(renderer)
            if m:
                msg = f"{what}: {m.decode()}"
        raise WfError(status, msg)
