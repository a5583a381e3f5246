"""Batched inverse rendering: recover a base-color texture (and the emission
strengths of the light quads) from many views of one scene.

Every view is a scene slot of one batch, so a forward pass renders all views
in one set of launches and the path-replay adjoint of that frame accumulates
the texel and emission gradients of every view with atomics into one texture
gradient.  Paths are recomputed by the adjoint, never stored.  The images are
compared with target images rendered from the true texture; Adam updates the
texture (torch), clamped to [TEXEL_MIN, 1].

Path replay obtains a texel's gradient as the remaining radiance of a path
vertex divided by the vertex's albedo; at an albedo of exactly 0 both are 0
and the gradient is reported as 0.  The optimiser therefore keeps texels at
or above TEXEL_MIN, where the gradient is the derivative of the (nearly
linear) estimate and a texel driven towards 0 can come back.

`python -m wavefrontpathtracing_amd.inverse --steps 200` runs the full
optimisation on the 256-instance scene (64 orbit views at 512x512, 8 spp,
depth 5, 1024x1024 texture, 4 emissive quads).
"""
from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np

from . import _lib as L
from . import scenes
from .renderer import Material, Renderer, camera_create

TEXEL_MIN = 1e-4   # lower clamp of optimised texels (see the module docstring)


def orbit_views(n: int, radius: float, center, seed: int):
    """n camera positions spread over a sphere (golden-angle spiral, rotated by
    a seeded angle), all looking at `center`, upper hemisphere only."""
    rng = scenes.Rng(seed, 0x5eed0030)
    phase = 2.0 * math.pi * rng.float()
    out = []
    for k in range(n):
        z = 0.15 + 0.8 * (k + 0.5) / n                  # elevation: sin in [0.15, 0.95]
        r = math.sqrt(max(0.0, 1.0 - z * z))
        a = phase + k * math.pi * (3.0 - math.sqrt(5.0))
        eye = (center[0] + radius * r * math.cos(a), center[1] + radius * z, center[2] + radius * r * math.sin(a))
        out.append((eye, tuple(center)))
    return out


class InverseProblem:
    """All views of `scene` as one batch on one renderer; texture 0 of the
    scene is the unknown.  The TLASes use the union-area metric by default
    (inverse rendering has no reference counterpart whose tree it must
    keep; the union tree gives the same hits in fewer node visits)."""

    def __init__(self, scene: scenes.Scene, views, width: int, height: int, spp: int, max_bounces: int,
                 vfov_deg: float = 60.0, tlas_metric: str = "union"):
        if not scene.textures:
            raise ValueError("the scene needs a base-color texture to optimise")
        self.scene, self.views, self.width, self.height = scene, views, width, height
        self.r = Renderer(width, height)
        r = self.r
        r.set_tlas_metric(tlas_metric)
        mesh_h = [r.create_render_mesh(v, i) for v, i in scene.meshes]
        tw, th, rgb = scene.textures[0]
        self.tex_shape = (th, tw)
        self.tex = r.create_render_texture(tw, th, rgb)
        r.begin_batch()
        for eye, target in views:
            r.batch_add_scene(camera_create(eye, target, vfov_deg), width, height)
            for ins in scene.instances:
                mat = Material(base_color_factor=(*ins.base_color, 1.0), emissive_factor=ins.emissive,
                               emissive_strength=ins.emissive_strength,
                               base_color_texture=self.tex if ins.texture >= 0 else L.INVALID_HANDLE)
                r.submit_render_mesh(mesh_h[ins.mesh], ins.transform, mat)
        r.end_scene()
        r.update_settings(max_bounces=max_bounces, spp=spp, next_event_estimation=1, hdr_env_strength=0.0,
                          frame_seed=0, rr_start_bounce=3, rr_survival=0.75, accumulate=0, use_hip_graph=1)
        n_inst = len(scene.instances)
        self.emitters = [k for k, ins in enumerate(scene.instances) if ins.emissive_strength > 0.0 and any(ins.emissive)]
        # instance index (all views) -> emitter number, or -1
        self.inst_emitter = np.full(n_inst * len(views), -1, dtype=np.int64)
        for v in range(len(views)):
            for j, k in enumerate(self.emitters):
                self.inst_emitter[v * n_inst + k] = j
        self.base_strength = np.array([ins.emissive_strength for ins in scene.instances] * len(views), np.float32)

    def strengths(self, torch, em):
        """Per-instance strengths for all views from the emitter scalars `em`."""
        idx = torch.as_tensor(self.inst_emitter)
        base = torch.as_tensor(self.base_strength)
        gathered = em[idx.clamp(min=0)]
        return torch.where(idx >= 0, gathered, base)

    def render(self, torch, texels, em):
        from . import autograd as WA
        return WA.render(self.r, texels, self.strengths(torch, em), self.tex)

    def close(self):
        self.r.exit()


def run(steps: int = 200, views: int = 64, res: int = 512, spp: int = 8, max_bounces: int = 5, tex_size: int = 1024,
        lr: float = 1e-2, seed: int = 7, log=print, scene: scenes.Scene = None, frame_seed: int = 0,
        fit_emission: bool = True, tlas_metric: str = "union"):
    """Optimise the texture (from a constant 0.5) and, with fit_emission, the
    light strengths (from half their value) towards the scene's own; returns
    per-step records (loss, seconds, rays).  frame_seed 0 draws a new frame seed every render,
    so the loss includes Monte Carlo noise; a fixed seed renders every step
    with the target's own random numbers."""
    import torch
    sc = scene if scene is not None else scenes.tlas_scene(texture=True)
    if scene is None:
        sc.textures = [(tex_size, tex_size, scenes.value_noise_texture(seed, tex_size, tex_size))]
    vs = orbit_views(views, 26.0, (0.0, 0.0, 0.0), seed)
    prob = InverseProblem(sc, vs, res, res, spp, max_bounces, tlas_metric=tlas_metric)
    prob.r.update_settings(frame_seed=frame_seed)
    dev = torch.device("cuda:0")
    try:
        th, tw = prob.tex_shape
        t_true = torch.ones((th, tw, 4), dtype=torch.float32, device=dev)
        t_true[..., :3] = torch.from_numpy(sc.textures[0][2]).to(dev)
        em_true = torch.tensor([sc.instances[k].emissive_strength for k in prob.emitters], dtype=torch.float32)
        with torch.no_grad():
            target = prob.render(torch, t_true, em_true)[:, :3].clone()
        texels = torch.full_like(t_true, 0.5)
        texels[..., 3] = 1.0
        texels.requires_grad_(True)
        em = (em_true * (0.5 if fit_emission else 1.0)).clone().requires_grad_(fit_emission)
        groups = [{"params": [texels], "lr": lr}]
        if fit_emission:   # the light strengths are O(10) where texels are O(1): ten times the step
            groups.append({"params": [em], "lr": 10.0 * lr})
        opt = torch.optim.Adam(groups)
        hist = []
        for it in range(steps):
            t0 = time.perf_counter()
            opt.zero_grad()
            img = prob.render(torch, texels, em)
            loss = ((img[:, :3] - target) ** 2).mean()
            loss.backward()
            with torch.no_grad():
                texels.grad[..., 3] = 0.0
            opt.step()
            with torch.no_grad():
                texels.clamp_(TEXEL_MIN, 1.0)
                if fit_emission:
                    em.clamp_(min=1e-3)
            torch.cuda.synchronize()
            st = prob.r.frame_stats()
            rec = {"step": it, "loss": float(loss.item()), "s": time.perf_counter() - t0,
                   "texel_mae": float((texels.detach()[..., :3] - t_true[..., :3]).abs().mean().item()),
                   "texel_min": float(texels.detach()[..., :3].min().item()),
                   "emission": [round(float(x), 4) for x in em.detach()], "forward_rays": st.rays_extended + st.shadow_rays}
            hist.append(rec)
            if log:
                log(json.dumps(rec))
        return hist
    finally:
        prob.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--views", type=int, default=64)
    ap.add_argument("--res", type=int, default=512)
    ap.add_argument("--spp", type=int, default=8)
    ap.add_argument("--depth", type=int, default=5)
    ap.add_argument("--tex", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=1e-2)
    a = ap.parse_args()
    run(a.steps, a.views, a.res, a.spp, a.depth, a.tex, a.lr)


if __name__ == "__main__":
    main()
