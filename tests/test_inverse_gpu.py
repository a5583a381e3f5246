"""Batched adjoint and inverse rendering on the GPU.

A batch of views renders every view with the path numbering it has alone,
so the adjoint of the whole batch equals the sum of the adjoints of the
views rendered one at a time (up to the order of float atomics).  A few Adam
steps on the texture lower the multi-view loss.
"""
import numpy as np
import pytest

from wavefrontpathtracing_amd import scenes
from wavefrontpathtracing_amd import inverse as INV

pytestmark = pytest.mark.gpu


def textured_small_scene():
    sc = scenes.small_tlas_scene()
    sc.textures = [(16, 16, scenes.value_noise_texture(3, 16, 16, cells=4))]
    for ins in sc.instances:
        if ins.emissive_strength == 0.0:
            ins.texture = 0
    return sc


def test_batch_adjoint_equals_sum_of_view_adjoints():
    sc = textured_small_scene()
    views = INV.orbit_views(3, 22.0, (0.0, 0.0, 0.0), 5)
    prob = INV.InverseProblem(sc, views, 40, 32, spp=2, max_bounces=3)
    r = prob.r
    try:
        r.update_settings(use_hip_graph=0, frame_seed=9)
        n_pix = r.image_pixels()
        rng = np.random.default_rng(1)
        dl = np.zeros((n_pix, 4), np.float32)
        dl[:, :3] = rng.normal(size=(n_pix, 3))
        n_inst = len(sc.instances) * len(views)
        r.render()
        g_all, e_all = r.render_adjoint(dl, prob.tex, prob.tex_shape, n_inst)
        g_sum = np.zeros_like(g_all)
        e_sum = np.zeros_like(e_all)
        for k in range(len(views)):
            r.render_scenes(k, 1)
            g, e = r.render_adjoint(dl, prob.tex, prob.tex_shape, n_inst)
            g_sum += g
            e_sum += e
    finally:
        prob.close()
    assert np.abs(g_all).max() > 0
    assert np.abs(g_all - g_sum).max() <= 1e-4 * np.abs(g_all).max()
    em = [v * len(sc.instances) + k for v in range(len(views)) for k in prob.emitters]
    assert np.allclose(e_all[em], e_sum[em], rtol=1e-4, atol=1e-6 * np.abs(e_all).max())


def test_inverse_steps_lower_the_loss():
    sc = textured_small_scene()
    hist = INV.run(steps=6, views=3, res=40, spp=2, max_bounces=3, lr=0.03, seed=4, log=None, scene=sc, frame_seed=11,
                   fit_emission=False)
    assert hist[-1]["loss"] < 0.7 * hist[0]["loss"], [h["loss"] for h in hist]
    assert hist[-1]["texel_mae"] < hist[0]["texel_mae"]


def test_texels_near_zero_keep_their_gradient():
    """Path replay divides a vertex's remaining radiance by its albedo, so
    where the bilinear albedo is exactly 0 (every texel of the footprint at
    0, here a 3x3 block around texel j in one channel) the texels get no
    gradient; inverse.run keeps optimised texels at or above TEXEL_MIN
    instead.  With the block at TEXEL_MIN the gradient of the image sum is
    the derivative of an estimate that is linear in the texels (up to paths
    that meet the block twice): equal within 1 % with the block at TEXEL_MIN
    and at 2 * TEXEL_MIN, and positive.  A run with a learning rate that
    would drive texels below 0 ends with every texel >= TEXEL_MIN."""
    import torch
    sc = textured_small_scene()
    views = INV.orbit_views(2, 22.0, (0.0, 0.0, 0.0), 5)
    prob = INV.InverseProblem(sc, views, 40, 32, spp=4, max_bounces=3)
    try:
        prob.r.update_settings(frame_seed=13)
        th, tw = prob.tex_shape
        em = torch.tensor([sc.instances[k].emissive_strength for k in prob.emitters], dtype=torch.float32)

        def grad_at(j, value):
            t = torch.full((th, tw, 4), 0.5, dtype=torch.float32, device="cuda")
            t[..., 3] = 1.0
            if j is not None:
                y, x = divmod(j, tw)
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        t[(y + dy) % th, (x + dx) % tw, 0] = value
            t.requires_grad_(True)
            prob.render(torch, t, em)[:, :3].sum().backward()
            return t.grad.view(-1, 4)[:, 0].detach().cpu().numpy()

        g = grad_at(None, None)
        j = int(np.argmax(g))
        assert g[j] > 0
        assert grad_at(j, 0.0)[j] == 0.0
        g1, g2 = grad_at(j, INV.TEXEL_MIN)[j], grad_at(j, 2 * INV.TEXEL_MIN)[j]
        assert g1 > 0 and abs(g1 - g2) <= 0.01 * g2, (g1, g2)
    finally:
        prob.close()
    hist = INV.run(steps=3, views=2, res=40, spp=2, max_bounces=3, lr=2.0, seed=4, log=None, scene=textured_small_scene(),
                   frame_seed=11, fit_emission=False)
    assert min(h["texel_min"] for h in hist) >= INV.TEXEL_MIN * (1 - 1e-6)
