"""Pins the FP64 reference gradient (path-replay adjoint over recorded paths)
against FP64 central finite differences (h = 1e-4) of the same estimator."""
import numpy as np
import pytest

from wavefrontpathtracing_amd import scenes

import oracle_api as O

H = 1e-4
REL_TOL = 1e-3


def small_diff_scene(emissive_texture=False):
    """The differentiable room; with `emissive_texture`, the light also
    carries a seeded-noise emissive texture (texture 1)."""
    sc = scenes.differentiable_scene(width=24, height=20, ico_subdiv=2, tex_size=16)
    if emissive_texture:
        sc.textures.append((8, 8, scenes.value_noise_texture(13, 8, 8, cells=4, lo=0.0, hi=2.0)))
        sc.instances[3].emissive_texture = 1
    return sc


@pytest.fixture(scope="module", params=[(0, False), (1, False), (0, True), (1, True)],
                ids=["bsdf", "nee", "bsdf-emtex", "nee-emtex"])
def diff(request):
    nee, emtex = request.param
    sc = small_diff_scene(emtex)
    orc = O.OracleScene(sc)
    return orc.record(max_bounces=4, spp=8, nee=nee, env=0.0, frame_seed=7)


def _loss_and_grad(img, target):
    r = img - target
    return float((r * r).sum()), 2.0 * r


def test_fp64_image_matches_fp32_render(diff):
    """Re-weighting the recorded paths with the scene's own values gives the
    FP32 render to FP32 rounding."""
    img64 = diff.image()
    orc = diff.orc
    img32 = orc.render(max_bounces=4, spp=8, nee=diff.prm.nee, env=0.0, frame_seed=7)[..., :3]
    assert np.abs(img64 - img32).max() <= 1e-5 * max(1.0, np.abs(img64).max())
    assert img64.sum() > 0


def test_texel_gradients_match_central_differences(diff):
    rng = np.random.default_rng(3)
    t0 = diff.texels0()
    target = diff.image(texels=np.clip(t0 + rng.normal(0, 0.1, t0.shape), 0.05, 0.95))
    _, dl = _loss_and_grad(diff.image(), target)
    g_tex, _ = diff.adjoint(dl)
    # 32 texel channels with non-negligible gradients
    mag = np.abs(g_tex[:, :3]).reshape(-1)
    cand = np.nonzero(mag > 1e-3 * mag.max())[0]
    pick = rng.choice(cand, size=min(32, len(cand)), replace=False)
    for flat in pick:
        j, c = divmod(int(flat), 3)
        tp, tm = t0.copy(), t0.copy()
        tp[j, c] += H
        tm[j, c] -= H
        fd = (_loss_and_grad(diff.image(texels=tp), target)[0] - _loss_and_grad(diff.image(texels=tm), target)[0]) / (2 * H)
        assert abs(fd - g_tex[j, c]) <= REL_TOL * abs(fd), (j, c, fd, g_tex[j, c])


def test_emission_gradient_matches_central_differences(diff):
    s0 = diff.strength0
    target = diff.image(strength=s0 * 0.8)
    _, dl = _loss_and_grad(diff.image(), target)
    _, g_s = diff.adjoint(dl)
    k = 3   # the ceiling light
    sp, sm = s0.copy(), s0.copy()
    sp[k] += H
    sm[k] -= H
    fd = (_loss_and_grad(diff.image(strength=sp), target)[0] - _loss_and_grad(diff.image(strength=sm), target)[0]) / (2 * H)
    assert abs(fd) > 0 and abs(fd - g_s[k]) <= REL_TOL * abs(fd), (fd, g_s[k])
    assert np.all(g_s[:3] == 0.0)   # non-emissive instances get no gradient


def test_texture_that_also_emits_is_refused():
    """Differentiating a texture that is also an emissive texture is refused:
    only its albedo use would be a variable."""
    sc = small_diff_scene()
    sc.instances[3].emissive_texture = 0
    d = O.OracleScene(sc).record(max_bounces=2, spp=1, nee=1, env=0.0, frame_seed=7)
    with pytest.raises(ValueError):
        d.image()
    with pytest.raises(ValueError):
        d.adjoint(np.zeros((sc.height, sc.width, 3)))
    d.image(tex=-1)
