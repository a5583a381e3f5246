"""Pins the CPU oracle's intersection, camera and traversal functions with
closed-form known answers, invariants and brute force (the reference ships no
tests or golden vectors for this path).

Each check is chosen so that a plausible restatement mistake fails it:
float vs double epsilon, strict vs non-strict comparisons, back-face culling
sign, minNum/maxNum NaN handling in the slab test, child order, and the
glm association order of matrix products.
"""
import ctypes as C

import numpy as np
import pytest

from wavefrontpathtracing_amd import scenes
from wavefrontpathtracing_amd import _lib as L
from wavefrontpathtracing_amd import renderer as R

import oracle_api as O

F = np.float32
FLT_MAX = F(3.4028235e38)


def tri(v0, v1, v2, o, d, t=FLT_MAX):
    a = [O.f32(x) for x in (v0, v1, v2, o, d)]
    tt = np.array([t], dtype=np.float32)
    bary = np.zeros(2, dtype=np.float32)
    hit = O.olib().orc_intersect_triangle(*[O.p(x) for x in a], O.p(tt), O.p(bary))
    return hit, tt[0], bary


def slab(bmin, bmax, o, d, t=FLT_MAX):
    d = O.f32(d)
    with np.errstate(divide="ignore"):
        inv = (F(1.0) / d).astype(np.float32)
    a = [O.f32(bmin), O.f32(bmax), O.f32(o), inv]
    return F(O.olib().orc_intersect_aabb(*[O.p(x) for x in a], float(t)))


# ---- Moller-Trumbore (intersect.hlsl:7-51) ---------------------------------------
V0, V1, V2 = (0, 0, 0), (1, 0, 0), (0, 1, 0)    # front face normal +z


def test_triangle_front_hit_exact_t_and_bary():
    hit, t, b = tri(V0, V1, V2, (0.25, 0.25, 1.0), (0, 0, -1))
    assert hit == 1 and t == F(1.0) and b[0] == F(0.25) and b[1] == F(0.25)


def test_triangle_back_face_culled():
    hit, t, _ = tri(V0, V1, V2, (0.25, 0.25, -1.0), (0, 0, 1))
    assert hit == 0 and t == FLT_MAX


def test_triangle_det_threshold_is_float_1e8():
    """det == float(1e-8) = 9.99999994e-09 is accepted (det < eps misses);
    a double epsilon 1e-8 would reject it."""
    x = F(1e-8)
    hit, _, _ = tri((0, 0, 0), (x, 0, 0), (0, 1, 0), (x / 4, 0.25, 1.0), (0, 0, -1))
    assert hit == 1
    x2 = np.nextafter(x, F(0))
    hit2, _, _ = tri((0, 0, 0), (x2, 0, 0), (0, 1, 0), (x2 / 4, 0.25, 1.0), (0, 0, -1))
    assert hit2 == 0


def test_triangle_edges_and_vertices_inclusive():
    # v = 0 (on edge v0-v2), w = 0, v + w == 1 and the vertex itself are inside
    for o, bary in [((0.0, 0.5, 1.0), (0.0, 0.5)), ((0.5, 0.0, 1.0), (0.5, 0.0)),
                    ((0.5, 0.5, 1.0), (0.5, 0.5)), ((1.0, 0.0, 1.0), (1.0, 0.0))]:
        hit, t, b = tri(V0, V1, V2, o, (0, 0, -1))
        assert hit == 1 and t == F(1.0) and tuple(b) == bary
    hit, _, _ = tri(V0, V1, V2, (np.nextafter(F(0), F(-1)), 0.5, 1.0), (0, 0, -1))
    assert hit == 0


def test_triangle_t_bounds():
    # t == ray.t is rejected (t >= ray.t), just below is accepted, t = 0 accepted, t < 0 rejected
    assert tri(V0, V1, V2, (0.25, 0.25, 1.0), (0, 0, -1), t=F(1.0))[0] == 0
    assert tri(V0, V1, V2, (0.25, 0.25, 1.0), (0, 0, -1), t=np.nextafter(F(1.0), F(2)))[0] == 1
    hit, t, _ = tri(V0, V1, V2, (0.25, 0.25, 0.0), (0, 0, -1))
    assert hit == 1 and t == F(0.0)
    assert tri(V0, V1, V2, (0.25, 0.25, -0.5), (0, 0, -1))[0] == 0


def test_triangle_matches_float64_moller_trumbore():
    """Random front-facing hits: t and barycentrics within a few ulps of the
    exact (float64) solution of o + t d = v0 + v e1 + w e2."""
    rng = np.random.default_rng(4)
    n_hit = 0
    for _ in range(400):
        v = rng.uniform(-1, 1, (3, 3)).astype(np.float32)
        o = rng.uniform(-3, 3, 3).astype(np.float32)
        tgt = v[0] + 0.3 * (v[1] - v[0]) + 0.3 * (v[2] - v[0])
        d = (tgt - o).astype(np.float32)
        hit, t, b = tri(v[0], v[1], v[2], o, d)
        e1, e2 = (v[1] - v[0]).astype(np.float64), (v[2] - v[0]).astype(np.float64)
        front = np.dot(np.cross(e1, e2), d.astype(np.float64)) < 0
        A = np.stack([-d.astype(np.float64), e1, e2], axis=1)
        tx, vx, wx = np.linalg.solve(A, (o - v[0]).astype(np.float64))
        if not front:
            assert hit == 0
            continue
        assert hit == 1
        n_hit += 1
        assert abs(t - tx) <= 1e-4 * max(1.0, abs(tx)) and abs(b[0] - vx) < 1e-4 and abs(b[1] - wx) < 1e-4
    assert n_hit > 100


# ---- slab test (intersect.hlsl:53-76) ----------------------------------------------
def test_slab_entry_distance_and_misses():
    assert slab((1, -1, -1), (2, 1, 1), (0, 0, 0), (1, 0, 0)) == F(1.0)
    assert slab((1, -1, -1), (2, 1, 1), (0, 0, 0), (-1, 0, 0)) == FLT_MAX       # behind
    assert slab((-1, -1, -1), (1, 1, 1), (0, 0, 0), (1, 0, 0)) == F(-1.0)       # inside: negative tmin
    assert slab((1, -1, -1), (2, 1, 1), (0, 0, 0), (1, 0, 0), t=F(1.0)) == FLT_MAX   # tmin < ray.t is strict
    assert slab((1, 2, -1), (2, 3, 1), (0, 0, 0), (1, 0, 0)) == FLT_MAX        # parallel, outside in y


def test_slab_nan_from_zero_times_inf():
    """Origin exactly on a slab plane with a zero direction component gives
    0 * inf = NaN.  minNum/maxNum drop the NaN and keep the other plane's
    +-inf, so a ray lying in a box face misses; a box that is flat in that
    axis (both planes through the origin) has NaN for both and the axis is
    ignored entirely, so the ray hits."""
    assert slab((0, -1, 1), (1, 1, 2), (0, 0, 0), (0, 0, 1)) == FLT_MAX
    assert slab((-1, -1, 1), (0, 1, 2), (0, 0, 0), (0, 0, 1)) == FLT_MAX
    assert slab((0, -1, 1), (0, 1, 2), (0, 0, 0), (0, 0, 1)) == F(1.0)


def test_slab_grazing_edge_counts_as_hit():
    # tmax == tmin (ray touching the box's (1,1) edge only) is a hit
    assert slab((1, 1, -1), (2, 2, 1), (0, 2, 0), (1, -1, 0)) == F(1.0)
    assert slab((1, 1, -1), (2, 2, 1), (0, np.nextafter(F(2), F(0)), 0), (1, -1, 0)) == FLT_MAX


# ---- glm (mat4 inverse / products / camera) ------------------------------------
def test_mat4_inverse_close_to_float64_and_bitwise_shared():
    rng = scenes.Rng(77)
    for _ in range(50):
        t = rng.transform(0.5, 2.0, 10.0)
        inv = np.zeros(16, np.float32)
        O.olib().orc_mat4_inverse(O.p(t), O.p(inv))
        ref = np.linalg.inv(t.reshape(4, 4).T.astype(np.float64)).T.reshape(16)
        assert np.allclose(inv, ref, rtol=1e-5, atol=1e-5)
        rec = R.make_bvh_instance(t, np.zeros(3), np.ones(3), 0)
        assert rec["world_to_local"].tobytes() == inv.tobytes()


def test_mat4_mul_vec4_association_is_glm_pairwise():
    """glm computes (c0*x + c1*y) + (c2*z + c3*w); choose values where the
    left-to-right sum rounds differently."""
    m = np.zeros(16, np.float32)
    m[0], m[4], m[8], m[12] = 1.0, 1e8, -1e8, 1.0   # row 0 coefficients
    v = np.array([1.0, 1.0, 1.0, 1.0], np.float32)
    out = np.zeros(4, np.float32)
    O.olib().orc_mat4_mul_vec4(O.p(m), O.p(v), O.p(out))
    assert out[0] == F(F(1.0 + 1e8) + F(-1e8 + 1.0))   # pairwise: 1e8 + (-1e8) = 0
    assert out[0] == F(0.0)


def test_camera_closed_form():
    v = O.camera_view((0, 0, -3), (0, 0, 0), 60.0, 512, 512)
    w2v = np.array(v.world_to_view).reshape(4, 4).T
    assert np.array_equal(w2v, np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 3], [0, 0, 0, 1]], np.float32))
    p = np.array(v.view_to_clip).reshape(4, 4).T
    h = F(np.cos(F(0.5) * F(60 * 0.017453292519943295)) / np.sin(F(0.5) * F(60 * 0.017453292519943295)))
    assert abs(p[1, 1] - 1.7320508) < 1e-6 and p[0, 0] == p[1, 1] and p[3, 2] == 1.0
    rays = O.primary_rays(v, 512, 512)
    d = rays["direction"]
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    c = rays[256 * 512 + 256]
    assert c["direction"][2] > 0.9999 and abs(c["origin"][2] + 3.0) < 1e-6
    # vertical field of view: top-row center ray is ~30 degrees above the axis
    top = rays[0 * 512 + 256]["direction"]
    assert abs(np.degrees(np.arctan2(top[1], top[2])) - 30.0 * (511.0 / 512.0)) < 0.05
    del h


def test_camera_matches_library_view():
    for eye, tgt, w, h in [((0, 0, -3), (0, 0, 0), 512, 512), ((0, 5, -30), (0, 0, 0), 1920, 1080), ((3, 1, 2), (-1, 0.5, 4), 640, 480)]:
        ov = O.camera_view(eye, tgt, 60.0, w, h)
        lv = R.make_view(R.camera_create(eye, tgt, 60.0), w, h)
        assert bytes(ov) == bytes(lv)


def test_library_primary_rays_match_oracle():
    """The library's host restatement of the generate stage's rays equals the
    oracle's bit for bit."""
    for eye, tgt, w, h in [((0, 0, -3), (0, 0, 0), 512, 512), ((0, 5, -30), (0, 0, 0), 320, 180)]:
        sc = scenes.Scene(meshes=[], instances=[], eye=eye, target=tgt, vfov_deg=60.0, width=w, height=h)
        lib_rays = scenes.primary_rays(sc)
        orc_rays = O.primary_rays(O.camera_view(eye, tgt, 60.0, w, h), w, h)
        assert lib_rays.tobytes() == orc_rays.tobytes()


# ---- traversal (accelstruct.hlsl:58-218) vs brute force -----------------------
@pytest.mark.parametrize("which", ["soup", "tlas"])
def test_traversal_equals_brute_force(which):
    if which == "soup":
        sc = scenes.soup_scene(3000, seed=5, width=48, height=48)
    else:
        sc = scenes.tlas_scene(seed=9, width=48, height=40, n_instances=24, hf_quads=10, ico_subdiv=2, soup_tris=400)
    orc = O.OracleScene(sc)
    view = O.camera_view(sc.eye, sc.target, sc.vfov_deg, sc.width, sc.height)
    bmin, bmax = orc.tnodes[0]["aabb_min"], orc.tnodes[0]["aabb_max"]
    rays = np.concatenate([O.primary_rays(view, sc.width, sc.height), scenes.sphere_rays(2, bmin, bmax, 1500)])
    bvh = orc.trace_closest(rays)
    brute = orc.brute_closest(rays)
    assert (bvh["instance_idx"] != L.MISS).sum() > 100
    assert np.array_equal(bvh["t"].view(np.uint32), brute["t"].view(np.uint32))
    same_id = (bvh["instance_idx"] == brute["instance_idx"]) & (bvh["primitive_idx"] == brute["primitive_idx"])
    # ids may only differ where two triangles tie at exactly the same t
    assert same_id.mean() > 0.999
    vis = orc.trace_any(rays)
    assert np.array_equal(vis == 1, bvh["instance_idx"] == L.MISS)


def test_traversal_counters_are_consistent():
    sc = scenes.soup_scene(2000, seed=6, width=32, height=32)
    orc = O.OracleScene(sc)
    view = O.camera_view(sc.eye, sc.target, sc.vfov_deg, 32, 32)
    _, cnt = orc.trace_closest(O.primary_rays(view, 32, 32), counters=True)
    # one TLAS leaf (root), one instance per ray that enters it; BLAS node
    # fetches are 1 (root) + 2 per interior node processed: always odd
    assert np.all(cnt["tlas_nodes"] == 1) and np.all(cnt["instances"] <= 1)
    entered = cnt["instances"] == 1
    assert np.all(cnt["blas_nodes"][entered] % 2 == 1)


def test_union_tlas_gives_the_same_hits_in_fewer_visits():
    """The union-metric TLAS changes only the order in which instances are
    visited: closest hits (ids, t, barycentrics) equal the reference TLAS's
    and brute force, and rays visit far fewer TLAS nodes (the reference's
    "max of mins / max of maxes" metric pairs boxes that are far apart)."""
    sc = scenes.tlas_scene(seed=9, width=48, height=40, n_instances=64, hf_quads=10, ico_subdiv=2, soup_tris=400)
    ref = O.OracleScene(sc)
    uni = O.OracleScene(sc, tlas_metric=O.TLAS_UNION)
    assert ref.tnodes.tobytes() != uni.tnodes.tobytes()
    bmin, bmax = ref.tnodes[0]["aabb_min"], ref.tnodes[0]["aabb_max"]
    rays = scenes.sphere_rays(3, bmin, bmax, 1 << 13)
    h_ref, c_ref = ref.trace_closest(rays, nthreads=8, counters=True)
    h_uni, c_uni = uni.trace_closest(rays, nthreads=8, counters=True)
    assert (h_ref["instance_idx"] != L.MISS).sum() > 1000
    assert len(O.hits_equal(h_ref, h_uni)) == 0
    brute = uni.brute_closest(rays[:2000])
    assert np.array_equal(h_uni["t"][:2000].view(np.uint32), brute["t"].view(np.uint32))
    assert c_uni["tlas_nodes"].mean() < 0.7 * c_ref["tlas_nodes"].mean(), (c_uni["tlas_nodes"].mean(), c_ref["tlas_nodes"].mean())
    assert np.array_equal(uni.trace_any(rays) == 1, h_uni["instance_idx"] == L.MISS)
