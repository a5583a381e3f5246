"""Procedural workloads (the benchmark's and the parity tests' inputs).

Each builder returns a ``Scene``: meshes as (wf_vertex array, index array),
instances as (mesh id, local_to_world, material), a camera and a resolution.
All randomness comes from the library's seeded PCG32, so a scene is the same
bytes on every host.

* ``soup_scene``   - one random-triangle soup seen from (0,0,-3) looking +z,
                     60 deg vertical field of view (65,536 triangles at 512^2
                     is the CPU-runnable closest-hit configuration);
* ``tlas_scene``   - 256 instances of 4 BLAS (value-noise heightfield
                     256x256 quads, icosphere subdiv 6, 200K-triangle soup,
                     12-quad box room) with uniform random rotations, scale
                     U(0.5, 2), translation U[-10,10]^3, Lambertian albedo
                     U(0.2, 0.8), plus 4 downward emissive quads (radiance 10)
                     above the scene; camera (0,5,-30) -> origin, 1920x1080.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from ._lib import VERTEX, check, lib, ptr


@dataclass
class Instance:
    mesh: int
    transform: np.ndarray            # (16,) float32 column-major local_to_world
    base_color: tuple = (0.8, 0.8, 0.8)
    emissive: tuple = (0.0, 0.0, 0.0)
    emissive_strength: float = 0.0
    texture: int = -1                # index into Scene.textures, -1 = none
    emissive_texture: int = -1       # index into Scene.textures, -1 = none (emission = factor * strength * texel)
    normal_texture: int = -1         # index into Scene.textures, -1 = none (read by the debug view modes only)
    metallic_roughness_texture: int = -1   # likewise


@dataclass
class Scene:
    meshes: list                     # [(vertices VERTEX[], indices u32[])]
    instances: list                  # [Instance]
    eye: tuple
    target: tuple
    vfov_deg: float
    width: int
    height: int
    textures: list = field(default_factory=list)   # [(w, h, rgb float32 (h, w, 3))]
    name: str = ""
    env_texture: int = -1            # equirect environment (index into textures), -1 = the constant default

    @property
    def triangle_counts(self):
        return [len(i) // 3 for _, i in self.meshes]

    def instanced_triangles(self) -> int:
        tc = self.triangle_counts
        return sum(tc[i.mesh] for i in self.instances)


# ---- mesh generators (thin wrappers of wf_scene.h) -----------------------------
def random_soup(seed: int, n: int, extent: float = 1.0, off_min: float = 0.005, off_max: float = 0.02):
    v = np.zeros(3 * n, dtype=VERTEX)
    i = np.zeros(3 * n, dtype=np.uint32)
    check(lib().wfs_random_soup(seed, n, extent, off_min, off_max, ptr(v), ptr(i)), what="random_soup")
    return v, i


def heightfield(seed: int, nx: int, nz: int, amplitude: float = 0.1):
    v = np.zeros((nx + 1) * (nz + 1), dtype=VERTEX)
    i = np.zeros(6 * nx * nz, dtype=np.uint32)
    check(lib().wfs_heightfield(seed, nx, nz, amplitude, ptr(v), ptr(i)), what="heightfield")
    return v, i


def icosphere(subdiv: int):
    nv, ni = C.c_uint32(), C.c_uint32()
    check(lib().wfs_icosphere(subdiv, None, None, C.byref(nv), C.byref(ni)), what="icosphere")
    v = np.zeros(nv.value, dtype=VERTEX)
    i = np.zeros(ni.value, dtype=np.uint32)
    check(lib().wfs_icosphere(subdiv, ptr(v), ptr(i), C.byref(nv), C.byref(ni)), what="icosphere")
    return v, i


def box_room():
    v = np.zeros(48, dtype=VERTEX)
    i = np.zeros(72, dtype=np.uint32)
    check(lib().wfs_box_room(ptr(v), ptr(i)), what="box_room")
    return v, i


def quad():
    v = np.zeros(4, dtype=VERTEX)
    i = np.zeros(6, dtype=np.uint32)
    check(lib().wfs_quad(ptr(v), ptr(i)), what="quad")
    return v, i


def value_noise_texture(seed: int, w: int, h: int, cells: int = 16, lo: float = 0.1, hi: float = 0.9):
    out = np.zeros((h, w, 3), dtype=np.float32)
    check(lib().wfs_value_noise_texture(seed, w, h, cells, lo, hi, ptr(out)), what="value_noise_texture")
    return out


def sphere_rays(seed: int, bmin, bmax, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=L.RAY)
    mn = np.ascontiguousarray(bmin, dtype=np.float32)
    mx = np.ascontiguousarray(bmax, dtype=np.float32)
    check(lib().wfs_sphere_rays(seed, ptr(mn), ptr(mx), n, ptr(out)), what="sphere_rays")
    return out


def primary_rays(scene: "Scene") -> np.ndarray:
    """The generate stage's camera rays for the scene's camera and size."""
    from .renderer import camera_create, make_view
    view = make_view(camera_create(scene.eye, scene.target, scene.vfov_deg), scene.width, scene.height)
    out = np.zeros(scene.width * scene.height, dtype=L.RAY)
    check(lib().wfs_primary_rays(C.byref(view), scene.width, scene.height, ptr(out)), what="primary_rays")
    return out


def bounce_rays(scene: "Scene", rays: np.ndarray, hits: np.ndarray, seed: int = 7) -> np.ndarray:
    """First-bounce rays of a path tracer: for every ray of `rays` that hit
    (closest hits `hits`), a cosine-distributed direction about the hit's
    interpolated world normal, from the hit point nudged by 1e-8 along the
    new direction (make_ray, common.hlsl:145-154).  Directions come from a
    seeded numpy generator, so the set is fixed for a scene, ray set and
    seed; it is a workload, not the path tracer's own random numbers."""
    hit = np.nonzero(hits["instance_idx"] != 0xFFFFFFFF)[0]
    inst, prim, bary = hits["instance_idx"][hit], hits["primitive_idx"][hit], hits["bary"][hit].astype(np.float64)
    o = rays["origin"][hit].astype(np.float64)
    d = rays["direction"][hit].astype(np.float64)
    P = o + hits["t"][hit, None].astype(np.float64) * d
    N = np.zeros_like(P)
    for k, ins in enumerate(scene.instances):
        sel = np.nonzero(inst == k)[0]
        if len(sel) == 0:
            continue
        v, i = scene.meshes[ins.mesh]
        tri = i.reshape(-1, 3)[prim[sel]]
        n0, n1, n2 = (v["normal"][tri[:, j]].astype(np.float64) for j in range(3))
        ln = n0 + bary[sel, :1] * (n1 - n0) + bary[sel, 1:] * (n2 - n0)
        M = np.asarray(ins.transform, np.float64).reshape(4, 4).T
        N[sel] = ln @ M[:3, :3].T
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    rng = np.random.default_rng(seed)
    r0, r1 = rng.random(len(hit)), rng.random(len(hit))
    a = np.where(np.abs(N[:, :1]) > np.abs(N[:, 2:]), np.stack([-N[:, 1], N[:, 0], 0 * N[:, 0]], 1),
                 np.stack([0 * N[:, 0], -N[:, 2], N[:, 1]], 1))
    t = a / np.linalg.norm(a, axis=1, keepdims=True)
    b = np.cross(N, t)
    st, ph = np.sqrt(r0), 2.0 * np.pi * r1
    L_ = (st * np.cos(ph))[:, None] * t + np.sqrt(1.0 - r0)[:, None] * N + (st * np.sin(ph))[:, None] * b
    L_ /= np.linalg.norm(L_, axis=1, keepdims=True)
    out = np.zeros(len(hit), dtype=L.RAY)
    out["origin"] = (P + L_ * 1e-8).astype(np.float32)
    out["direction"] = L_.astype(np.float32)
    out["tmin"] = np.float32(1e-8)
    out["tmax"] = np.finfo(np.float32).max
    return out


class Rng:
    def __init__(self, seed: int, seq: int = 0x5eed0010):
        self._s = L.Pcg32()
        lib().wfs_pcg32_seed(C.byref(self._s), seed, seq)

    def next(self) -> int:
        return lib().wfs_pcg32_next(C.byref(self._s))

    def float(self) -> float:
        return lib().wfs_pcg32_float(C.byref(self._s))

    def uniform(self, lo: float, hi: float) -> float:
        return float(np.float32(lo) + np.float32(hi - lo) * np.float32(self.float()))

    def transform(self, scale_min: float, scale_max: float, t_extent: float) -> np.ndarray:
        out = np.zeros(16, dtype=np.float32)
        lib().wfs_random_transform(C.byref(self._s), scale_min, scale_max, t_extent, ptr(out))
        return out


def identity() -> np.ndarray:
    return np.eye(4, dtype=np.float32).reshape(16)


def translate_scale(t, s) -> np.ndarray:
    m = np.eye(4, dtype=np.float32)
    m[0, 0], m[1, 1], m[2, 2] = s
    m[0:3, 3] = t
    return np.ascontiguousarray(m.T).reshape(16)   # column-major


# ---- workloads ---------------------------------------------------------------------
def soup_scene(n: int = 65536, seed: int = 1, width: int = 512, height: int = 512) -> Scene:
    v, i = random_soup(seed, n)
    return Scene(meshes=[(v, i)], instances=[Instance(0, identity())], eye=(0.0, 0.0, -3.0), target=(0.0, 0.0, 0.0),
                 vfov_deg=60.0, width=width, height=height, name=f"soup{n}")


def tlas_scene(seed: int = 2, width: int = 1920, height: int = 1080, n_instances: int = 256,
               hf_quads: int = 256, ico_subdiv: int = 6, soup_tris: int = 200_000, emitters: int = 4,
               texture: bool = False) -> Scene:
    """256 instances of 4 BLAS + emissive quads (see module docstring)."""
    meshes = [heightfield(seed, hf_quads, hf_quads, 0.1), icosphere(ico_subdiv), random_soup(seed + 100, soup_tris), box_room()]
    rng = Rng(seed)
    inst = []
    for k in range(n_instances):
        t = rng.transform(0.5, 2.0, 10.0)
        albedo = (rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8))
        inst.append(Instance(k % 4, t, base_color=albedo, texture=0 if texture else -1))
    textures = [(1024, 1024, value_noise_texture(seed + 7, 1024, 1024))] if texture else []
    if emitters:
        meshes.append(quad())
        q = len(meshes) - 1
        for e in range(emitters):
            x = -6.0 if e % 2 == 0 else 6.0
            z = -6.0 if e // 2 == 0 else 6.0
            inst.append(Instance(q, translate_scale((x, 14.0, z), (6.0, 1.0, 6.0)), base_color=(0.0, 0.0, 0.0),
                                 emissive=(1.0, 1.0, 1.0), emissive_strength=10.0))
    return Scene(meshes=meshes, instances=inst, eye=(0.0, 5.0, -30.0), target=(0.0, 0.0, 0.0), vfov_deg=60.0,
                 width=width, height=height, textures=textures, name="tlas256")


def small_tlas_scene(seed: int = 12, width: int = 160, height: int = 120) -> Scene:
    """A reduced instance of tlas_scene for fast parity tests."""
    return tlas_scene(seed=seed, width=width, height=height, n_instances=64, hf_quads=24, ico_subdiv=3,
                      soup_tris=3000, emitters=4)


def differentiable_scene(seed: int = 5, width: int = 512, height: int = 512, ico_subdiv: int = 6, tex_size: int = 512,
                         tex_seed: int = None) -> Scene:
    """Box room (12 quads, half-size 2) with one emissive ceiling quad and two
    icospheres; every diffuse surface takes its albedo from one RGB texture of
    periodic value noise in [0.1, 0.9] (texture 0).  The emission scalar is
    the ceiling quad's emissive strength (instance 3)."""
    tex = value_noise_texture(seed if tex_seed is None else tex_seed, tex_size, tex_size, cells=16, lo=0.1, hi=0.9)
    meshes = [box_room(), icosphere(ico_subdiv), quad()]
    inst = [Instance(0, translate_scale((0.0, 0.0, 0.0), (2.0, 2.0, 2.0)), base_color=(1.0, 1.0, 1.0), texture=0),
            Instance(1, translate_scale((-0.8, -1.35, 0.3), (0.6, 0.6, 0.6)), base_color=(1.0, 1.0, 1.0), texture=0),
            Instance(1, translate_scale((0.9, -1.45, -0.4), (0.55, 0.55, 0.55)), base_color=(1.0, 1.0, 1.0), texture=0),
            Instance(2, translate_scale((0.0, 1.95, 0.0), (1.2, 1.0, 1.2)), base_color=(0.0, 0.0, 0.0),
                     emissive=(1.0, 1.0, 1.0), emissive_strength=8.0)]
    return Scene(meshes=meshes, instances=inst, eye=(0.0, 0.2, -1.9), target=(0.0, -0.6, 0.5), vfov_deg=70.0,
                 width=width, height=height, textures=[(tex_size, tex_size, tex)], name="diff")


def batch_scenes(seed: int = 4, n_scenes: int = 64, tri_min: int = 10_000, tri_max: int = 2_000_000,
                 resolutions=(256, 512, 1024)):
    """Uneven scenes for the batched renderer: each a random-triangle soup
    (log-uniform triangle count in [tri_min, tri_max]) inside a 12-quad box
    room of half-size 1.5 lit by one emissive ceiling quad, a square
    resolution drawn from `resolutions`, and its own camera inside the room."""
    rng = Rng(seed, 0x5eed0020)
    room_v, room_i = box_room()
    quad_v, quad_i = quad()
    out = []
    lo, hi = np.log(tri_min), np.log(tri_max)
    for s in range(n_scenes):
        n = int(np.exp(lo + (hi - lo) * rng.float()))
        res = resolutions[rng.next() % len(resolutions)]
        ang = 2.0 * np.pi * rng.float()
        eye = (1.35 * np.sin(ang), 0.2 + 0.4 * rng.float(), 1.35 * np.cos(ang))
        soup = random_soup(seed * 1000 + s, n)
        meshes = [soup, (room_v, room_i), (quad_v, quad_i)]
        inst = [Instance(0, identity(), base_color=(rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8))),
                Instance(1, translate_scale((0.0, 0.0, 0.0), (1.5, 1.5, 1.5)), base_color=(0.7, 0.7, 0.7)),
                Instance(2, translate_scale((0.0, 1.45, 0.0), (1.2, 1.0, 1.2)), base_color=(0.0, 0.0, 0.0),
                         emissive=(1.0, 1.0, 1.0), emissive_strength=6.0)]
        out.append(Scene(meshes=meshes, instances=inst, eye=eye, target=(0.0, 0.0, 0.0), vfov_deg=70.0,
                         width=res, height=res, name=f"batch{s}"))
    return out


def load_batch(renderer, scene_list):
    """begin_batch, then per scene: create its meshes (the room and light
    meshes are shared when they are the same arrays), open a slot and submit."""
    from .renderer import Material, camera_create
    cache = {}
    handles = []
    for sc in scene_list:
        hs = []
        for v, i in sc.meshes:
            key = (v.ctypes.data, len(v), i.ctypes.data, len(i))
            if key not in cache:
                cache[key] = renderer.create_render_mesh(v, i)
            hs.append(cache[key])
        handles.append(hs)
    renderer.begin_batch()
    for sc, hs in zip(scene_list, handles):
        renderer.batch_add_scene(camera_create(sc.eye, sc.target, sc.vfov_deg), sc.width, sc.height)
        for ins in sc.instances:
            mat = Material(base_color_factor=(*ins.base_color, 1.0), emissive_factor=ins.emissive,
                           emissive_strength=ins.emissive_strength)
            renderer.submit_render_mesh(hs[ins.mesh], ins.transform, mat)
    renderer.end_scene()


def load(renderer, scene: Scene, begin: bool = True):
    """Create meshes/textures and submit instances (the caller side of
    renderer::create_render_mesh / submit_render_mesh, core/scene.cpp:92-105)."""
    from .renderer import Material, camera_create
    mesh_h = [renderer.create_render_mesh(v, i) for v, i in scene.meshes]
    tex_h = [renderer.create_render_texture(w, h, rgb) for (w, h, rgb) in scene.textures]
    if begin:
        env = tex_h[scene.env_texture] if scene.env_texture >= 0 else L.INVALID_HANDLE
        renderer.begin_scene(camera_create(scene.eye, scene.target, scene.vfov_deg), env)
    for ins in scene.instances:
        mat = Material(base_color_factor=(*ins.base_color, 1.0), emissive_factor=ins.emissive,
                       emissive_strength=ins.emissive_strength,
                       base_color_texture=tex_h[ins.texture] if ins.texture >= 0 else L.INVALID_HANDLE,
                       emissive_texture=tex_h[ins.emissive_texture] if ins.emissive_texture >= 0 else L.INVALID_HANDLE,
                       normal_texture=tex_h[ins.normal_texture] if ins.normal_texture >= 0 else L.INVALID_HANDLE,
                       metallic_roughness_texture=(tex_h[ins.metallic_roughness_texture]
                                                   if ins.metallic_roughness_texture >= 0 else L.INVALID_HANDLE))
        renderer.submit_render_mesh(mesh_h[ins.mesh], ins.transform, mat)
    return mesh_h, tex_h
