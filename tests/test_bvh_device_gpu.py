"""Device BLAS builder (wf_bvh_build_device): the blob equals the oracle's
restatement of bvh_builder.cpp byte for byte, so nodes come out in the
reference's allocation order with the reference's triangle index order.

Sizes straddle the 64-triangle boundary between the level passes and the
one-wave subtree builder; the special meshes cover zero centroid extents,
ties between bins, signed zeros in the bounds (whose sign follows the first
zero in the reference's scan order), denormal extents and every supported
interval count."""
import numpy as np
import pytest

from wavefrontpathtracing_amd import renderer as R
from wavefrontpathtracing_amd import scenes
from wavefrontpathtracing_amd import _lib as L

import oracle_api as O

pytestmark = pytest.mark.gpu


def _soup(n, seed):
    return O.deindex(*scenes.random_soup(seed, n))


def _signed_zero_soup(n, seed, frac=0.3):
    """A soup where a fraction of the coordinates are +0.0 or -0.0, so many node
    bounds are zeros whose sign depends on the order the builder scans."""
    pos = _soup(n, seed).reshape(-1, 3, 3).copy()
    rng = np.random.default_rng(seed)
    m = rng.random(pos.shape) < frac
    neg = rng.random(pos.shape) < 0.5
    pos[m] = np.where(neg[m], np.float32(-0.0), np.float32(0.0))
    return pos.reshape(-1, 9)


def _integer_soup(n, seed):
    """Vertices on a coarse integer lattice: many equal centroids and bin ties."""
    rng = np.random.default_rng(seed)
    c = rng.integers(-4, 5, size=(n, 1, 3)).astype(np.float32)
    d = rng.integers(-1, 2, size=(n, 3, 3)).astype(np.float32)
    return (c + d).reshape(-1, 9)


def _denormal_chain(n=240):
    pos = np.zeros((n, 9), np.float32)
    for k in range(n):
        x = np.float32(2.0) ** np.float32(-k * 0.5)
        s = x * np.float32(0.01)
        pos[k] = [x, -s, 0, x, s, s, x, s, -s]
    return pos


@pytest.mark.parametrize("bins", [8, 16])
@pytest.mark.parametrize("n", [1, 2, 3, 17, 63, 64, 65, 100, 1024, 4097, 65536])
def test_device_blob_equals_oracle(n, bins):
    pos9 = _soup(n, seed=7 * n + bins)
    assert R.bvh_build_device(pos9, bins) == O.blas_blob(pos9, bins)


def _special(mesh):
    if mesh == "heightfield":
        return O.deindex(*scenes.heightfield(5, 40, 30, 0.1))
    if mesh == "icosphere":
        return O.deindex(*scenes.icosphere(4))
    if mesh == "box_room":
        return O.deindex(*scenes.box_room())
    if mesh == "coplanar_signed_zero":
        pos = _soup(3000, 3).reshape(-1, 3, 3).copy()
        pos[:, :, 2] = np.where(np.arange(9000).reshape(3000, 3) % 3 == 1, np.float32(-0.0), np.float32(0.0))
        return pos.reshape(-1, 9)
    if mesh == "duplicates":
        return np.tile(_soup(1, 4), (300, 1))
    if mesh == "signed_zero_small":
        return _signed_zero_soup(50, 11)
    if mesh == "signed_zero":
        return _signed_zero_soup(20000, 12)
    if mesh == "integer":
        return _integer_soup(5000, 13)
    if mesh == "denormal":
        return _denormal_chain()
    raise AssertionError(mesh)


@pytest.mark.parametrize("mesh", ["heightfield", "icosphere", "box_room", "coplanar_signed_zero", "duplicates",
                                  "signed_zero_small", "signed_zero", "integer", "denormal"])
def test_device_special_meshes(mesh):
    pos9 = _special(mesh)
    for bins in (2, 8, 16, 64):
        assert R.bvh_build_device(pos9, bins) == O.blas_blob(pos9, bins), f"{mesh}, {bins} bins"


@pytest.mark.parametrize("mesh", ["soup", "integer", "signed_zero", "denormal"])
def test_device_single_prim_mode(mesh):
    pos9 = {"soup": lambda: _soup(3000, 9), "integer": lambda: _integer_soup(2000, 5),
            "signed_zero": lambda: _signed_zero_soup(2000, 6), "denormal": _denormal_chain}[mesh]()
    for bins in (8, 16):
        assert R.bvh_build_device(pos9, bins, subdivide_single_prim=True) == O.blas_blob(pos9, bins, single_prim=True)


def test_device_large_soup_equals_host_builder():
    """1M-triangle soup at 16 intervals: the device blob equals the host builder's
    (the host builder equals the oracle on every pinned case)."""
    pos9 = _soup(1 << 20, 21)
    assert R.bvh_build_device(pos9, 16) == R.bvh_build(pos9, 16)


def test_device_build_rejects_bad_input():
    pos9 = _soup(100, 1)
    for bins in (1, 65):
        with pytest.raises(L.WfError):
            R.bvh_build_device(pos9, bins)
    bad = pos9.copy()
    bad[40, 4] = np.inf
    with pytest.raises(L.WfError):
        R.bvh_build_device(bad, 8)
    with pytest.raises(L.WfError):
        R.bvh_build_device(np.zeros((0, 9), np.float32), 8)


def test_renderer_device_and_host_builds_trace_alike(gpu_renderer_factory):
    """The 256-instance scene loaded once with every BLAS built on the host and
    once on the GPU: the renderers return the same closest hits bit for bit on
    camera rays and incoherent rays, and the same frame image."""
    sc = scenes.tlas_scene(width=160, height=90)
    outs = []
    for builder in ("host", "device"):
        r = gpu_renderer_factory(sc.width, sc.height)
        r.set_blas_builder(builder)
        scenes.load(r, sc)
        r.build_tlas()
        box = np.float32(12.0)   # instances are translated within [-10, 10]^3
        rays = np.concatenate([scenes.primary_rays(sc), scenes.sphere_rays(7, -box * np.ones(3, np.float32),
                                                                           box * np.ones(3, np.float32), 1 << 16)])
        hits = r.trace_closest(rays)
        r.update_settings(max_bounces=3, spp=1, next_event_estimation=1, hdr_env_strength=0.5, frame_seed=3)
        r.render()
        outs.append((hits, r.read_energy()))
    assert outs[0][0].tobytes() == outs[1][0].tobytes()
    assert np.array_equal(outs[0][1], outs[1][1])
