"""Batched rendering: many uneven scenes in one set of launches (optionally
one captured hipGraph).  A scene renders the same image alone or in the
batch, and each scene's image matches the oracle with the scene's path
numbering (mean relative per-pixel error <= 1e-3, image_metrics).  The
tests run with the extend kernel's large queue claims both as the queues'
capacity selects them (not at these sizes) and forced on, since the large
batch frames run with them."""
import numpy as np
import pytest

from wavefrontpathtracing_amd import Renderer, scenes, tuning

import oracle_api as O
from image_metrics import assert_image_close

pytestmark = pytest.mark.gpu

SETTINGS = dict(max_bounces=4, spp=2, next_event_estimation=1, hdr_env_strength=0.0, frame_seed=5, rr_start_bounce=3,
                rr_survival=0.75)


@pytest.fixture(scope="module")
def batch():
    sl = scenes.batch_scenes(seed=31, n_scenes=6, tri_min=500, tri_max=20_000, resolutions=(24, 40, 56))
    r = Renderer(64, 64)
    scenes.load_batch(r, sl)
    yield sl, r
    r.exit()


@pytest.fixture(params=[(0, 1), (1, 1), (0, 0), (1, 0)],
                ids=["claims_by_capacity", "big_claims", "claims_by_capacity_no_skip", "big_claims_no_skip"])
def claims(request, batch):
    _, r = batch
    big, skip = request.param
    tuning.set_extend_config(r, big_claims=big, skip_popped=skip)
    yield request.param
    tuning.set_extend_config(r)


def _path_bases(sl, spp):
    out, at = [], 0
    for sc in sl:
        out.append(at)
        at += sc.width * sc.height * spp
    return out


def test_batch_equals_per_scene_loop(batch, claims):
    sl, r = batch
    r.update_settings(use_hip_graph=0, **SETTINGS)
    r.render_scenes(0, len(sl))
    together = [r.read_scene_energy(k, sc.width, sc.height) for k, sc in enumerate(sl)]
    for k, sc in enumerate(sl):
        r.render_scenes(k, 1)
        alone = r.read_scene_energy(k, sc.width, sc.height)
        assert np.array_equal(alone, together[k]), f"scene {k}"


def test_big_claims_change_no_image(batch):
    """The claim size decides only which lane traces which ray."""
    sl, r = batch
    r.update_settings(use_hip_graph=0, **SETTINGS)
    imgs = []
    try:
        for big in (2, 1):
            tuning.set_extend_config(r, big_claims=big)
            r.render()
            imgs.append([r.read_scene_energy(k, sc.width, sc.height) for k, sc in enumerate(sl)])
    finally:
        tuning.set_extend_config(r)
    for k in range(len(sl)):
        assert np.array_equal(imgs[0][k], imgs[1][k]), f"scene {k}"


def test_batch_graph_replay_equals_direct_launches(batch, claims):
    sl, r = batch
    r.update_settings(use_hip_graph=0, **SETTINGS)
    r.render()
    direct = [r.read_scene_energy(k, sc.width, sc.height) for k, sc in enumerate(sl)]
    r.update_settings(use_hip_graph=1, **SETTINGS)
    for _ in range(3):   # first frame captures, the others replay
        r.render()
        graph = [r.read_scene_energy(k, sc.width, sc.height) for k, sc in enumerate(sl)]
        for k in range(len(sl)):
            assert np.array_equal(direct[k], graph[k]), f"scene {k}"


def test_batch_scenes_match_oracle(batch, claims):
    sl, r = batch
    r.update_settings(use_hip_graph=1, **SETTINGS)
    r.render()
    bases = _path_bases(sl, SETTINGS["spp"])
    for k, sc in enumerate(sl):
        gpu = r.read_scene_energy(k, sc.width, sc.height)
        ref = O.OracleScene(sc).render(max_bounces=4, spp=2, nee=1, env=0.0, frame_seed=5, path_base=bases[k], nthreads=8)
        assert_image_close(gpu, ref, f"scene {k}")
        assert gpu[..., :3].sum() > 0
