"""Python mirror of the reference renderer front-end (source/renderer/renderer.h:13-55).

``Renderer`` wraps one ``wf_renderer`` handle of libwf.so; method names and
argument meaning follow ``renderer::init / exit / begin_frame / end_frame /
begin_scene / render / end_scene / create_render_texture /
create_render_mesh / submit_render_mesh``.  Errors the reference would
assert on raise ``WfError`` with the library's status and message.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from ._lib import HIT, RAY, VERTEX, WfError, check, lib, ptr

TEXTURE_FORMAT_RGBA8 = 1
TEXTURE_FORMAT_RGBA8_SRGB = 2
TEXTURE_FORMAT_RGB32F = 3
TEXTURE_FORMAT_RGBA32F = 4
TEXTURE_FORMAT_RG8 = 5


@dataclass
class Material:
    """material_asset_t (core/assets/asset_types.h:10-21)."""
    base_color_factor: tuple = (1.0, 1.0, 1.0, 1.0)
    metallic_factor: float = 0.0
    roughness_factor: float = 1.0
    emissive_factor: tuple = (0.0, 0.0, 0.0)
    emissive_strength: float = 0.0
    base_color_texture: int = L.INVALID_HANDLE
    normal_texture: int = L.INVALID_HANDLE
    metallic_roughness_texture: int = L.INVALID_HANDLE
    emissive_texture: int = L.INVALID_HANDLE

    def to_c(self) -> L.MaterialAsset:
        m = L.MaterialAsset()
        m.base_color_factor[:] = [float(x) for x in self.base_color_factor]
        m.metallic_factor = self.metallic_factor
        m.roughness_factor = self.roughness_factor
        m.emissive_factor[:] = [float(x) for x in self.emissive_factor]
        m.emissive_strength = self.emissive_strength
        m.base_color_texture = self.base_color_texture
        m.normal_texture = self.normal_texture
        m.metallic_roughness_texture = self.metallic_roughness_texture
        m.emissive_texture = self.emissive_texture
        return m


def camera_create(view_origin, target_pos, vfov_deg: float) -> L.Camera:
    """camera::create (core/camera/camera.cpp:8-16)."""
    cam = L.Camera()
    eye = np.ascontiguousarray(view_origin, dtype=np.float32)
    tgt = np.ascontiguousarray(target_pos, dtype=np.float32)
    check(lib().wf_camera_create(C.byref(cam), ptr(eye), ptr(tgt), float(vfov_deg)), what="camera_create")
    return cam


def make_view(camera: L.Camera, width: int, height: int) -> L.View:
    """The view_t renderer::begin_scene derives from a camera (renderer.cpp:599-630)."""
    v = L.View()
    check(lib().wf_make_view(C.byref(camera), width, height, C.byref(v)), what="make_view")
    return v


class Renderer:
    def __init__(self, render_width: int, render_height: int, backbuffer_count: int = 2, vsync: bool = False):
        """renderer::init (renderer.cpp:222-482)."""
        p = L.InitParams(render_width, render_height, backbuffer_count, int(vsync))
        h = C.c_void_p()
        check(lib().wf_renderer_init(C.byref(p), C.byref(h)), what="renderer init")
        self._h = h
        self.render_width = render_width
        self.render_height = render_height
        self._batch_dims = None

    # ---- lifecycle ----
    @property
    def handle(self):
        if self._h is None:
            raise WfError(6, "renderer has exited")
        return self._h

    def exit(self) -> None:
        if self._h is not None:
            lib().wf_renderer_exit(self._h)
            self._h = None

    def __del__(self):
        try:
            self.exit()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.exit()

    def _check(self, status, what):
        check(status, self.handle, what)

    @property
    def stream(self) -> int:
        """Address of the hipStream_t the renderer orders its GPU work on."""
        s = C.c_void_p()
        self._check(lib().wf_renderer_get_stream(self.handle, C.byref(s)), "get_stream")
        return s.value

    def begin_frame(self):
        self._check(lib().wf_renderer_begin_frame(self.handle), "begin_frame")

    def end_frame(self):
        self._check(lib().wf_renderer_end_frame(self.handle), "end_frame")

    def begin_scene(self, camera: L.Camera, env_render_texture_handle: int = L.INVALID_HANDLE):
        self._check(lib().wf_renderer_begin_scene(self.handle, C.byref(camera), env_render_texture_handle), "begin_scene")
        self._batch_dims = None

    def render(self):
        self._check(lib().wf_renderer_render(self.handle), "render")

    def end_scene(self):
        self._check(lib().wf_renderer_end_scene(self.handle), "end_scene")

    # ---- batches (many scenes per frame) ----
    def begin_batch(self):
        self._check(lib().wf_renderer_begin_batch(self.handle), "begin_batch")
        self._batch_dims = []

    def batch_add_scene(self, camera: L.Camera, width: int, height: int) -> int:
        out = C.c_uint32()
        self._check(lib().wf_renderer_batch_add_scene(self.handle, C.byref(camera), width, height, C.byref(out)), "batch_add_scene")
        self._batch_dims.append((width, height))
        return out.value

    def image_pixels(self) -> int:
        """Pixels of one frame's image: every batch scene's, concatenated, or
        the render resolution outside a batch."""
        if self._batch_dims is not None:
            return sum(w * h for w, h in self._batch_dims)
        return self.render_width * self.render_height

    def scene_count(self) -> int:
        out = C.c_uint32()
        self._check(lib().wf_renderer_get_scene_count(self.handle, C.byref(out)), "get_scene_count")
        return out.value

    def render_scenes(self, first: int, count: int):
        self._check(lib().wf_renderer_render_scenes(self.handle, first, count), "render_scenes")

    def read_scene_energy(self, scene: int, width: int, height: int) -> np.ndarray:
        out = np.empty((height, width, 4), dtype=np.float32)
        self._check(lib().wf_renderer_read_scene_energy(self.handle, scene, ptr(out)), "read_scene_energy")
        return out

    # ---- differentiable rendering ----
    def render_adjoint(self, dl_dimage: np.ndarray, texture: int, texture_shape, n_instances: int):
        """Path-replay adjoint of the last render(): returns (dLoss/dtexels
        (h, w, 4), dLoss/d emission strength per instance)."""
        dl = np.ascontiguousarray(dl_dimage, dtype=np.float32)
        if texture != L.INVALID_HANDLE:
            th, tw = texture_shape
            g_tex = np.zeros((th, tw, 4), dtype=np.float32)
            gp = ptr(g_tex)
        else:
            g_tex, gp = None, None
        g_em = np.zeros(n_instances, dtype=np.float32)
        self._check(lib().wf_renderer_render_adjoint(self.handle, ptr(dl), texture, gp, ptr(g_em)), "render_adjoint")
        return g_tex, g_em

    def render_adjoint_device(self, d_dl: int, texture: int, d_grad_tex: int, d_grad_em: int) -> None:
        self._check(lib().wf_renderer_render_adjoint_device(self.handle, C.c_void_p(d_dl), texture, C.c_void_p(d_grad_tex),
                                                            C.c_void_p(d_grad_em)), "render_adjoint_device")

    def update_texture(self, texture: int, rgba: np.ndarray) -> None:
        a = np.ascontiguousarray(rgba, dtype=np.float32)
        self._check(lib().wf_renderer_update_texture(self.handle, texture, ptr(a)), "update_texture")

    def update_texture_device(self, texture: int, d_rgba: int) -> None:
        self._check(lib().wf_renderer_update_texture_device(self.handle, texture, C.c_void_p(d_rgba)), "update_texture_device")

    def set_emission_strength(self, instance: int, strength: float) -> None:
        self._check(lib().wf_renderer_set_emission_strength(self.handle, instance, float(strength)), "set_emission_strength")

    def read_energy_device(self, d_out: int) -> None:
        self._check(lib().wf_renderer_read_energy_device(self.handle, C.c_void_p(d_out)), "read_energy_device")

    # ---- resources ----
    def create_render_texture(self, width: int, height: int, data: np.ndarray, fmt: int = TEXTURE_FORMAT_RGB32F) -> int:
        data = np.ascontiguousarray(data)
        bpp = {TEXTURE_FORMAT_RGBA8: 32, TEXTURE_FORMAT_RGBA8_SRGB: 32, TEXTURE_FORMAT_RGB32F: 96, TEXTURE_FORMAT_RGBA32F: 128,
               TEXTURE_FORMAT_RG8: 16}[fmt]
        need = width * height * bpp // 8
        if data.nbytes != need:
            raise WfError(1, f"texture data holds {data.nbytes} bytes, expected {need}")
        p = L.RenderTextureParams(width, height, bpp, fmt, ptr(data))
        out = C.c_uint64()
        self._check(lib().wf_renderer_create_render_texture(self.handle, C.byref(p), C.byref(out)), "create_render_texture")
        return out.value

    def create_render_mesh(self, vertices: np.ndarray, indices: np.ndarray) -> int:
        v = np.ascontiguousarray(vertices, dtype=VERTEX)
        i = np.ascontiguousarray(indices, dtype=np.uint32)
        p = L.RenderMeshParams(len(v), ptr(v), len(i), ptr(i))
        out = C.c_uint64()
        self._check(lib().wf_renderer_create_render_mesh(self.handle, C.byref(p), C.byref(out)), "create_render_mesh")
        return out.value

    def submit_render_mesh(self, render_mesh_handle: int, transform, material: Material) -> None:
        t = np.ascontiguousarray(transform, dtype=np.float32).reshape(16)
        m = material.to_c()
        self._check(lib().wf_renderer_submit_render_mesh(self.handle, render_mesh_handle, ptr(t), C.byref(m)), "submit_render_mesh")

    # ---- settings ----
    @property
    def settings(self) -> L.RenderSettings:
        s = L.RenderSettings()
        self._check(lib().wf_renderer_get_settings(self.handle, C.byref(s)), "get_settings")
        return s

    @settings.setter
    def settings(self, s: L.RenderSettings) -> None:
        self._check(lib().wf_renderer_set_settings(self.handle, C.byref(s)), "set_settings")

    def update_settings(self, **kw) -> None:
        s = self.settings
        for k, v in kw.items():
            setattr(s, k, v)
        self.settings = s

    def set_build_options(self, interval_count: int = 8, subdivide_single_prim: bool = False) -> None:
        o = L.BuildOptions(interval_count, int(subdivide_single_prim))
        self._check(lib().wf_renderer_set_build_options(self.handle, C.byref(o)), "set_build_options")

    BLAS_BUILDERS = {"auto": 0, "host": 1, "device": 2}

    def set_blas_builder(self, builder: str) -> None:
        """Where create_render_mesh builds BLASes: "auto" (the GPU for meshes of at
        least 16,384 triangles), "host" or "device"; all give the same nodes."""
        self._check(lib().wf_renderer_set_blas_builder(self.handle, self.BLAS_BUILDERS[builder]), "set_blas_builder")

    def set_tlas_builder(self, builder: str) -> None:
        """Where each scene's TLAS is built: "auto" (the GPU from 4,096 instances),
        "host" or "device"; all give the same nodes."""
        self._check(lib().wf_renderer_set_tlas_builder(self.handle, self.BLAS_BUILDERS[builder]), "set_tlas_builder")

    TLAS_METRICS = {"reference": 0, "union": 1}

    def set_tlas_metric(self, metric: str) -> None:
        """The match metric of the agglomerative TLAS builds: "reference" (the
        reference's tree, default) or "union" (half area of the union box; far
        fewer node visits, the same closest hits except exact ties between
        instances)."""
        self._check(lib().wf_renderer_set_tlas_metric(self.handle, self.TLAS_METRICS[metric]), "set_tlas_metric")

    def view(self) -> L.View:
        v = L.View()
        self._check(lib().wf_renderer_get_view(self.handle, C.byref(v)), "get_view")
        return v

    # ---- extend on caller rays ----
    def build_tlas(self) -> None:
        self._check(lib().wf_renderer_build_tlas(self.handle), "build_tlas")

    def scene_bounds(self):
        mn, mx = np.zeros(3, np.float32), np.zeros(3, np.float32)
        self._check(lib().wf_renderer_get_scene_bounds(self.handle, ptr(mn), ptr(mx)), "get_scene_bounds")
        return mn, mx

    def trace_closest(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=RAY)
        hits = np.empty(len(rays), dtype=HIT)
        self._check(lib().wf_trace_closest(self.handle, ptr(rays), len(rays), ptr(hits)), "trace_closest")
        return hits

    def trace_closest_counters(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, dtype=RAY)
        hits = np.empty(len(rays), dtype=HIT)
        cnt = np.empty((len(rays), 4), dtype=np.uint32)
        self._check(lib().wf_trace_closest_counters(self.handle, ptr(rays), len(rays), ptr(hits), ptr(cnt)), "trace_closest_counters")
        return hits, cnt

    def trace_occluded(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=RAY)
        vis = np.empty(len(rays), dtype=np.uint32)
        self._check(lib().wf_trace_occluded(self.handle, ptr(rays), len(rays), ptr(vis)), "trace_occluded")
        return vis

    def trace_closest_device(self, d_rays: int, n: int, d_hits: int) -> None:
        """Device pointers (e.g. torch tensor data_ptr()); ordered on the renderer's stream."""
        self._check(lib().wf_trace_closest_device(self.handle, C.c_void_p(d_rays), n, C.c_void_p(d_hits)), "trace_closest_device")

    def read_stack_overflows(self) -> int:
        """Rays that overflowed the traversal stack in the last trace call or frame (synchronises)."""
        out = C.c_uint32()
        self._check(lib().wf_renderer_read_stack_overflows(self.handle, C.byref(out)), "read_stack_overflows")
        return out.value

    # ---- readbacks ----
    def read_energy(self) -> np.ndarray:
        out = np.empty((self.render_height, self.render_width, 4), dtype=np.float32)
        self._check(lib().wf_renderer_read_energy(self.handle, ptr(out)), "read_energy")
        return out

    def read_accumulator(self) -> np.ndarray:
        out = np.empty((self.render_height, self.render_width, 4), dtype=np.float32)
        self._check(lib().wf_renderer_read_accumulator(self.handle, ptr(out)), "read_accumulator")
        return out

    def read_final_color(self) -> np.ndarray:
        out = np.empty((self.render_height, self.render_width, 4), dtype=np.uint8)
        self._check(lib().wf_renderer_read_final_color(self.handle, ptr(out)), "read_final_color")
        return out

    def frame_stats(self) -> L.FrameStats:
        s = L.FrameStats()
        self._check(lib().wf_renderer_get_frame_stats(self.handle, C.byref(s)), "frame_stats")
        return s


# ---- host builders (bvh_builder_t / tlas_builder_t) ----------------------------
def bvh_build(tri_positions: np.ndarray, interval_count: int = 8, subdivide_single_prim: bool = False) -> bytes:
    """bvh_builder_t::build + extract: returns the GPU blob (header ‖ nodes ‖ triangles ‖ indices)."""
    pos = np.ascontiguousarray(tri_positions, dtype=np.float32).reshape(-1, 9)
    opt = L.BuildOptions(interval_count, int(subdivide_single_prim))
    size = C.c_uint64()
    nodes = C.c_uint32()
    check(lib().wf_bvh_build(ptr(pos), len(pos), C.byref(opt), None, 0, C.byref(size), C.byref(nodes)), what="bvh_build")
    out = np.empty(size.value, dtype=np.uint8)
    check(lib().wf_bvh_build(ptr(pos), len(pos), C.byref(opt), ptr(out), size.value, C.byref(size), C.byref(nodes)), what="bvh_build")
    return out.tobytes()


def bvh_build_device(tri_positions: np.ndarray, interval_count: int = 8, subdivide_single_prim: bool = False) -> bytes:
    """The blob of bvh_build with the tree built on the GPU (wf_bvh_build_device): the
    positions are uploaded, the node and index arrays read back, and the header and
    triangles (original order) assembled as bvh_builder_t::extract does."""
    import struct
    import torch
    pos = np.ascontiguousarray(tri_positions, dtype=np.float32).reshape(-1, 9)
    n = len(pos)
    d_pos = torch.from_numpy(pos.copy()).to("cuda") if n else torch.empty((1, 9), dtype=torch.float32, device="cuda")
    d_nodes = torch.empty((max(2 * n, 2), 8), dtype=torch.int32, device="cuda")
    d_idx = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
    opt = L.BuildOptions(interval_count, int(subdivide_single_prim))
    count = C.c_uint32()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(lib().wf_bvh_build_device(C.c_void_p(d_pos.data_ptr()), n, C.byref(opt), C.c_void_p(d_nodes.data_ptr()),
                                    2 * n, C.c_void_p(d_idx.data_ptr()), C.byref(count), stream), what="bvh_build_device")
    nodes = d_nodes[:count.value].cpu().numpy().tobytes()
    idx = d_idx[:n].cpu().numpy().tobytes()
    k = count.value
    return struct.pack("<III", 12, 12 + 32 * k, 12 + 32 * k + 36 * n) + nodes + pos.tobytes() + idx


class DeviceBvhBuilder:
    """wf_bvh_builder: device BLAS builds (bvh_builder_t on the GPU) whose scratch is
    kept across builds.  Arguments are device addresses (torch data_ptr())."""

    def __init__(self):
        h = C.c_void_p()
        check(lib().wf_bvh_builder_create(C.byref(h)), what="bvh_builder_create")
        self.handle = h

    def build(self, d_pos: int, n: int, d_nodes: int, node_capacity: int, d_idx: int, interval_count: int = 8,
              subdivide_single_prim: bool = False, stream: int = 0) -> int:
        opt = L.BuildOptions(interval_count, int(subdivide_single_prim))
        count = C.c_uint32()
        s = lib().wf_bvh_builder_build_device(self.handle, C.c_void_p(d_pos), n, C.byref(opt), C.c_void_p(d_nodes),
                                              node_capacity, C.c_void_p(d_idx), C.byref(count), C.c_void_p(stream))
        if s != 0:
            raise WfError(s, f"bvh_builder_build_device: {lib().wf_bvh_builder_last_error(self.handle).decode()}")
        return count.value

    def last_build(self):
        """(levels built a level at a time, subtrees built one per wave, largest leaf)"""
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().wf_bvh_builder_last_build(self.handle, C.byref(a), C.byref(b), C.byref(c)), what="bvh_builder_last_build")
        return a.value, b.value, c.value

    def close(self) -> None:
        if self.handle:
            lib().wf_bvh_builder_destroy(self.handle)
            self.handle = None


def tlas_build(instances: np.ndarray, metric: str = "reference") -> bytes:
    """TLAS blob (wf_tlas_build; wf_tlas_build_metric for the "union" metric)."""
    inst = np.ascontiguousarray(instances, dtype=L.BVH_INSTANCE)
    size = C.c_uint64()
    nodes = C.c_uint32()
    mt = Renderer.TLAS_METRICS[metric]

    def call(out, cap):
        if metric == "reference":
            return lib().wf_tlas_build(ptr(inst), len(inst), out, cap, C.byref(size), C.byref(nodes))
        return lib().wf_tlas_build_metric(ptr(inst), len(inst), mt, out, cap, C.byref(size), C.byref(nodes))

    check(call(None, 0), what="tlas_build")
    out = np.empty(size.value, dtype=np.uint8)
    check(call(ptr(out), size.value), what="tlas_build")
    return out.tobytes()


def tlas_build_device(instances: np.ndarray, metric: str = "reference") -> bytes:
    """The blob of tlas_build with the TLAS built on the GPU (wf_tlas_build_device,
    wf_tlas_build_device_metric for the "union" metric)."""
    import struct
    import torch
    inst = np.ascontiguousarray(instances, dtype=L.BVH_INSTANCE)
    m = len(inst)
    d_inst = torch.from_numpy(inst.view(np.uint8).copy()).to("cuda") if m else torch.empty(1, dtype=torch.uint8, device="cuda")
    d_nodes = torch.empty((max(2 * m, 2), 8), dtype=torch.int32, device="cuda")
    count = C.c_uint32()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if metric == "reference":
        check(lib().wf_tlas_build_device(C.c_void_p(d_inst.data_ptr()), m, C.c_void_p(d_nodes.data_ptr()), 2 * m,
                                         C.byref(count), stream), what="tlas_build_device")
    else:
        check(lib().wf_tlas_build_device_metric(C.c_void_p(d_inst.data_ptr()), m, Renderer.TLAS_METRICS[metric],
                                                C.c_void_p(d_nodes.data_ptr()), 2 * m, C.byref(count), stream),
              what="tlas_build_device")
    nodes = d_nodes[:count.value].cpu().numpy().tobytes()
    return struct.pack("<II", 8, 8 + 32 * count.value) + nodes + inst.tobytes()


def make_bvh_instance(transform, blas_min, blas_max, bvh_index: int) -> np.ndarray:
    t = np.ascontiguousarray(transform, dtype=np.float32).reshape(16)
    mn = np.ascontiguousarray(blas_min, dtype=np.float32)
    mx = np.ascontiguousarray(blas_max, dtype=np.float32)
    out = np.zeros(1, dtype=L.BVH_INSTANCE)
    check(lib().wf_make_bvh_instance(ptr(t), ptr(mn), ptr(mx), bvh_index, ptr(out)), what="make_bvh_instance")
    return out[0]


def parse_bvh_blob(blob: bytes):
    """Split a BLAS blob into (nodes, triangles, indices) numpy views."""
    h = np.frombuffer(blob, dtype=np.uint32, count=3)
    n_nodes = (h[1] - h[0]) // 32
    n_tris = (h[2] - h[1]) // 36
    nodes = np.frombuffer(blob, dtype=L.BVH_NODE, count=n_nodes, offset=int(h[0]))
    tris = np.frombuffer(blob, dtype=L.BVH_TRIANGLE, count=n_tris, offset=int(h[1]))
    idx = np.frombuffer(blob, dtype=np.uint32, count=n_tris, offset=int(h[2]))
    return nodes, tris, idx


def parse_tlas_blob(blob: bytes):
    h = np.frombuffer(blob, dtype=np.uint32, count=2)
    n_nodes = (h[1] - h[0]) // 32
    n_inst = (len(blob) - h[1]) // 92
    nodes = np.frombuffer(blob, dtype=L.TLAS_NODE, count=n_nodes, offset=int(h[0]))
    inst = np.frombuffer(blob, dtype=L.BVH_INSTANCE, count=n_inst, offset=int(h[1]))
    return nodes, inst


def write_pfm(path: str, rgba: np.ndarray) -> None:
    """Linear (h, w, 4) float image -> PFM (RGB, little endian, bottom row first)."""
    a = np.ascontiguousarray(rgba, dtype=np.float32)
    h, w = a.shape[:2]
    check(lib().wf_write_pfm(str(path).encode(), ptr(a), w, h), what="write_pfm")


def write_ppm(path: str, rgba8: np.ndarray) -> None:
    """(h, w, 4) uint8 image -> binary PPM (RGB)."""
    a = np.ascontiguousarray(rgba8, dtype=np.uint8)
    h, w = a.shape[:2]
    check(lib().wf_write_ppm(str(path).encode(), ptr(a), w, h), what="write_ppm")
