"""Writes tests/golden/fixtures.json: SHA-256 digests (and sizes) of the
oracle's outputs at fixed seeds, for the builders and the traversal at the
benchmark workloads' sizes.  These are regression fixtures of the oracle (the
reference has no tests or golden vectors of its own); tests check the oracle,
the host builders and the GPU path against them.

Run from the repository root: python tests/golden/make_fixtures.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from wavefrontpathtracing_amd import scenes  # noqa: E402
import oracle_api as O  # noqa: E402


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def blas_cases():
    """(name, pos9, bins): the 65,536-triangle soup at 8 and 16 intervals and small soups."""
    out = []
    for n, bins in ((1, 8), (2, 8), (3, 8), (17, 8), (1024, 8), (1024, 16), (65536, 8), (65536, 16)):
        v, i = scenes.random_soup(1, n)
        out.append((f"soup_{n}_{bins}", np.ascontiguousarray(v["position"][i].reshape(-1, 9), dtype=np.float32), bins))
    return out


def tlas_instances(m: int):
    rng = scenes.Rng(m)
    bmin, bmax = np.array([-1, -0.5, -1], np.float32), np.array([1, 0.5, 1], np.float32)
    return np.array([O.instance_record(rng.transform(0.5, 2.0, 10.0), bmin, bmax, k % 4) for k in range(m)])


def traversal_cases():
    """(name, scene, rays): the soup's 512x512 camera rays (16 intervals) and
    the 256-instance scene's 2^21 incoherent rays (the extend benchmark)."""
    soup = scenes.soup_scene(65536, seed=1)
    tl = scenes.tlas_scene()
    return [("soup65k_camera", soup, 16, None), ("tlas256_incoherent", tl, 8, 1 << 21)]


def main():
    fx = {"blas": {}, "tlas": {}, "trace": {}}
    for name, pos9, bins in blas_cases():
        nodes, idx = O.blas_build(pos9, bins)
        fx["blas"][name] = {"n": len(pos9), "bins": bins, "node_count": len(nodes), "nodes_sha256": sha(nodes),
                            "indices_sha256": sha(idx)}
    for m in (1, 2, 3, 17, 256):
        inst = tlas_instances(m)
        for metric, key in ((O.TLAS_REFERENCE, "reference"), (O.TLAS_UNION, "union")):
            nodes = O.tlas_build(inst, metric)
            fx["tlas"][f"m{m}_{key}"] = {"m": m, "node_count": len(nodes), "nodes_sha256": sha(nodes)}
    for name, sc, bins, n_inc in traversal_cases():
        orc = O.OracleScene(sc, bins=bins)
        if n_inc is None:
            view = O.camera_view(sc.eye, sc.target, sc.vfov_deg, sc.width, sc.height)
            rays = O.primary_rays(view, sc.width, sc.height)
        else:
            bmin = np.min(orc.instances["aabb_min"], axis=0)
            bmax = np.max(orc.instances["aabb_max"], axis=0)
            rays = scenes.sphere_rays(3, bmin, bmax, n_inc)
        hits = orc.trace_closest(rays, nthreads=8)
        fx["trace"][name] = {"rays": len(rays), "rays_sha256": sha(rays), "hits_sha256": sha(hits),
                             "hit_fraction": round(float((hits["instance_idx"] != 0xFFFFFFFF).mean()), 6)}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.json")
    with open(path, "w") as f:
        json.dump(fx, f, indent=1, sort_keys=True)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
