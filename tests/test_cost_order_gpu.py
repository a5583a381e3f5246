"""The cost-ordered queue of the trace calls on caller rays.

The trace calls may claim the rays predicted to traverse longest first (the
traversal configuration's cost_order): a pre-pass puts each ray in one of 16
cost buckets (its chord through the scene bounds, weighted by its distance
from their centre) and lists the ray indices bucket by bucket, keeping the
callers' order within a bucket.  Checked here:
  * the order is a permutation of the rays, bucket-major and stable within a
    bucket, for ragged ray counts on both sides of 2^21 and for rays that miss
    the bounds, have a zero or NaN direction or an empty interval;
  * the buckets agree with a float64 restatement of the predictor except at
    bucket boundaries (the pre-pass uses the hardware reciprocal and square
    root estimates: a scheduling heuristic, not a result);
  * traces with the order on and off give hits bit-identical to the oracle's,
    through the device and host-array entry points, at sizes where the order
    is used (more than two rays per persistent lane).
"""
import ctypes as C

import numpy as np
import pytest

from wavefrontpathtracing_amd import scenes, tuning
from wavefrontpathtracing_amd import _lib as L
from wavefrontpathtracing_amd.renderer import Renderer

import oracle_api as O

pytestmark = pytest.mark.gpu


def _bucket_f64(rays, bmin, bmax):
    """The pre-pass's predictor in float64 (bucket 0 = longest)."""
    o = rays["origin"].astype(np.float64)
    d = rays["direction"].astype(np.float64)
    bmin = np.asarray(bmin, np.float64)
    bmax = np.asarray(bmax, np.float64)
    ln = np.sqrt((d * d).sum(1))
    with np.errstate(all="ignore"):
        inv = 1.0 / d
        ta, tb = (bmin - o) * inv, (bmax - o) * inv
        t0 = np.fmax(np.fmax(rays["tmin"].astype(np.float64), 0.0), np.fmin(ta, tb).max(1))
        t1 = np.fmin(rays["tmax"].astype(np.float64), np.fmax(ta, tb).min(1))
        seg = (t1 - t0) * ln
        c = 0.5 * (bmin + bmax)
        w = o - c
        proj = (w * d).sum(1) / np.where(ln > 0, ln, np.inf)
        dist2 = np.maximum((w * w).sum(1) - proj * proj, 0.0)
        diag2 = ((bmax - bmin) ** 2).sum()
        x = 16.0 * seg / np.sqrt(diag2) / (1.0 + 16.0 * dist2 / diag2)
    q = np.where(~(x > 0), 0, np.where(x >= 15, 15, np.floor(np.nan_to_num(x, nan=0.0, posinf=15.0))))
    return (15 - q).astype(np.int64)


def _special_rays(bmin, bmax, n):
    """Rays that miss the bounds, zero and NaN directions, empty intervals."""
    rng = np.random.default_rng(5)
    out = np.zeros(n, dtype=L.RAY)
    out["origin"] = np.asarray(bmax, np.float32) + rng.uniform(1, 5, (n, 3)).astype(np.float32)
    out["direction"] = rng.normal(size=(n, 3)).astype(np.float32)
    out["tmax"] = np.float32(3.4e38)
    out["direction"][0::4] = 0.0
    out["direction"][1::4, 1] = np.nan
    out["tmax"][2::4] = 0.0
    return out


@pytest.mark.parametrize("n", [100003, 3000017])
def test_cost_order_is_a_stable_bucket_sort(n):
    """100,003 rays: the scan of the per-block histograms (16 words per 4,096
    rays) is staged through LDS; 3,000,017 rays (more than 2^21): it runs on
    the histograms in global memory."""
    import torch
    sc = scenes.small_tlas_scene()
    orc = O.OracleScene(sc)
    bmin, bmax = orc.tnodes[0]["aabb_min"], orc.tnodes[0]["aabb_max"]
    rays = scenes.sphere_rays(23, bmin, bmax, n)
    rays[:400] = _special_rays(bmin, bmax, 400)
    d_rays = torch.from_numpy(np.frombuffer(rays.tobytes(), np.uint8).copy()).cuda()
    d_order = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    d_bucket = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    mn = (C.c_float * 3)(*map(float, bmin))
    mx = (C.c_float * 3)(*map(float, bmax))
    torch.cuda.synchronize()
    L.check(L.lib().wf_selftest_cost_order(C.c_void_p(d_rays.data_ptr()), n, mn, mx, C.c_void_p(d_order.data_ptr()),
                                           C.c_void_p(d_bucket.data_ptr())), None, "selftest_cost_order")
    order = d_order.cpu().numpy().astype(np.int64)
    bucket = d_bucket.cpu().numpy().astype(np.int64)
    assert np.array_equal(np.sort(order), np.arange(n)), "the order is not a permutation of the rays"
    assert bucket.min() >= 0 and bucket.max() <= 15
    b = bucket[order]
    assert np.all(np.diff(b) >= 0), "buckets are not claimed longest first"
    same = b[1:] == b[:-1]
    assert np.all(np.diff(order)[same] > 0), "the callers' order is not kept within a bucket"
    # every bucket in use on these rays, and the predictor's own values
    assert len(np.unique(bucket)) >= 8
    ref = _bucket_f64(rays, bmin, bmax)
    assert np.mean(ref == bucket) > 0.99
    assert np.all(np.abs(ref - bucket) <= 1)
    # rays with a NaN direction or an empty interval go last
    k = np.arange(400) % 4
    assert np.all(bucket[:400][(k == 1) | (k == 2)] == 15)


@pytest.mark.parametrize("block", [256, 64])
def test_cost_ordered_traces_equal_the_oracle(block):
    import torch
    sc = scenes.small_tlas_scene()
    r = Renderer(sc.width, sc.height)
    try:
        scenes.load(r, sc)
        r.build_tlas()
        orc = O.OracleScene(sc)
        bmin, bmax = orc.tnodes[0]["aabb_min"], orc.tnodes[0]["aabb_max"]
        n = 1200001   # above two rays per persistent lane of any configuration
        rays = scenes.sphere_rays(29, bmin, bmax, n)
        rays[:400] = _special_rays(bmin, bmax, 400)
        ref = orc.trace_closest(rays, nthreads=16)
        d_rays = torch.from_numpy(np.frombuffer(rays.tobytes(), np.uint8).copy()).cuda()
        d_hits = torch.empty((n + 100) * ref.dtype.itemsize, dtype=torch.uint8, device="cuda")
        for order in (1, 2):
            tuning.set_extend_config(r, block=block, cost_order=order)
            assert tuning.get_extend_config(r)["cost_order"] == order
            for m in (n, 700001):
                d_hits.fill_(0xAB)
                torch.cuda.synchronize()
                r.trace_closest_device(d_rays.data_ptr(), m, d_hits.data_ptr())
                torch.cuda.synchronize()
                raw = d_hits.cpu().numpy().tobytes()
                got = np.frombuffer(raw[:m * ref.dtype.itemsize], ref.dtype)
                bad = O.hits_equal(got, ref[:m])
                assert len(bad) == 0, f"cost_order {order}, {m} rays: {len(bad)} hits differ"
                assert np.all(np.frombuffer(raw[m * ref.dtype.itemsize:], np.uint8) == 0xAB)
            bad = O.hits_equal(r.trace_closest(rays), ref)
            assert len(bad) == 0, f"host-array entry point, cost_order {order}: {len(bad)} hits differ"
    finally:
        r.exit()


def test_cost_ordered_occlusion_and_counters_equal_the_oracle():
    """The other trace calls on caller rays at a size where the order is
    used: occlusion (any hit) and the counting trace, whose per-ray counters
    are written at each ray's own index like its hit."""
    sc = scenes.small_tlas_scene()
    r = Renderer(sc.width, sc.height)
    try:
        scenes.load(r, sc)
        r.build_tlas()
        orc = O.OracleScene(sc)
        bmin, bmax = orc.tnodes[0]["aabb_min"], orc.tnodes[0]["aabb_max"]
        n = 1100003
        rays = scenes.sphere_rays(31, bmin, bmax, n)
        rays["tmax"][::3] = np.float32(0.5)   # some occlusion queries end early
        want_vis = orc.trace_any(rays, nthreads=16)
        want_hits, want_cnt = orc.trace_closest(rays, nthreads=16, counters=True)
        for order in (1, 2):
            tuning.set_extend_config(r, cost_order=order)
            vis = r.trace_occluded(rays)
            assert np.array_equal(vis, want_vis), f"cost_order {order}: {(vis != want_vis).sum()} occlusion results differ"
            hits, cnt = r.trace_closest_counters(rays)
            assert len(O.hits_equal(hits, want_hits)) == 0
            got = cnt.view(np.uint32).reshape(n, 4)
            want = want_cnt.view(np.uint32).reshape(n, 4)
            assert np.array_equal(got, want), f"cost_order {order}: counters differ for {(got != want).any(1).sum()} rays"
    finally:
        r.exit()


def test_cost_order_setting_is_checked():
    sc = scenes.small_tlas_scene()
    r = Renderer(sc.width, sc.height)
    try:
        from wavefrontpathtracing_amd.renderer import WfError
        with pytest.raises(WfError):
            tuning.set_extend_config(r, cost_order=3)
        tuning.set_extend_config(r, cost_order=0)
        assert tuning.get_extend_config(r)["cost_order"] == 1   # the default is on
    finally:
        r.exit()
