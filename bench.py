"""Benchmark: incoherent closest-hit throughput of the extend stage.

Workload: the 256-instance TLAS scene over 4 BLAS (value-noise heightfield
256x256 quads, icosphere subdiv 6, 200K-triangle soup, 12-quad box room) plus
4 emissive quads, 413,016 unique / ~26M instanced triangles, camera
(0,5,-30), 1920x1080.  Rays: 2^21 incoherent rays (origin uniform in the
scene bounds, direction uniform on the sphere, seed 3).  One step = one
closest-hit pass over all rays, inputs resident in HBM.

Prints one JSON line (driver contract) with a roofline object for the extend
kernel (algorithmic bytes from the reference algorithm's per-ray traversal
counters, measured with the library's counting kernel) and the CPU oracle
timed on the host cores on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "incoherent closest-hit Mrays/s @1M tris, batched vs plain; adjoint spp/s"


def bytes_per_ray(counters: np.ndarray) -> float:
    """B_ray = 32 (ray) + 20 (hit) + 32*(TLAS + BLAS node fetches) + 92*instances
    + 40*triangle tests (36-byte triangle + 4-byte index), per-ray means."""
    c = counters.astype(np.float64).mean(axis=0)
    return 32.0 + 20.0 + 32.0 * (c[0] + c[1]) + 92.0 * c[2] + 40.0 * c[3]


def cpu_baseline(scene, rays, seconds_target=12.0):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_api as O
    orc = O.OracleScene(scene)
    threads = min(os.cpu_count() or 1, 16)
    n = 4096
    while True:
        sample = rays[:n]
        t0 = time.perf_counter()
        orc.trace_closest(sample, nthreads=threads)
        dt = time.perf_counter() - t0
        if dt > 1.0 or n >= len(rays):
            break
        n = min(len(rays), n * 4)
    # scale to ~seconds_target of work
    n_final = int(min(len(rays), max(n, n * seconds_target / max(dt, 1e-6))))
    sample = rays[:n_final]
    t0 = time.perf_counter()
    orc.trace_closest(sample, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(n_final / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"first {n_final} of the {len(rays)} incoherent rays, oracle traversal (C, -O3, no FP contraction), "
                      f"{threads} threads, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rays", type=int, default=1 << 21)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the path tracer is a single-GPU program (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from wavefrontpathtracing_amd import Renderer, scenes, tuning
    from wavefrontpathtracing_amd import _lib as L

    t0 = time.perf_counter()
    sc = scenes.tlas_scene()
    r = Renderer(sc.width, sc.height)
    scenes.load(r, sc)
    r.build_tlas()
    t_build = time.perf_counter() - t0
    bmin, bmax = r.scene_bounds()
    rays = scenes.sphere_rays(3, bmin, bmax, args.rays)

    dev = torch.device("cuda:0")
    d_rays = torch.from_numpy(rays.view(np.uint8)).to(dev)
    d_hits = torch.empty(len(rays) * 20, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    stream = torch.cuda.ExternalStream(r.stream)

    def step():
        r.trace_closest_device(d_rays.data_ptr(), len(rays), d_hits.data_ptr())

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    t_wall = time.perf_counter()
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall
    ms_total = ev0.elapsed_time(ev1)
    ms_step = ms_total / args.steps
    mrays = len(rays) / (ms_step * 1e-3) / 1e6

    hits = d_hits.cpu().numpy().view(L.HIT)
    hit_frac = float((hits["instance_idx"] != L.MISS).mean())
    _, counters = r.trace_closest_counters(rays)
    b_ray = bytes_per_ray(counters)
    achieved_gbs = mrays * 1e6 * b_ray / 1e9
    cfg = tuning.get_extend_config(r)

    out = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "256-instance TLAS over 4 BLAS + 4 emissive quads (413,016 unique / "
                        f"{sc.instanced_triangles():,} instanced tris), 2^21 incoherent rays (uniform origin in scene bounds, "
                        "uniform sphere directions), closest hit",
            "rays": len(rays),
            "hit_fraction": round(hit_frac, 4),
            "traversal": cfg,
            "scene_build_s": round(t_build, 2),
            "wall_s_timed": round(t_wall, 4),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved_gbs / HBM_PEAK_GBS, 4),
            "traffic": None,
            "bytes_per_ray": round(b_ray, 1),
            "mean_counters": {k: round(float(v), 3) for k, v in
                              zip(("tlas_nodes", "blas_nodes", "instances", "tri_tests"), counters.mean(axis=0))},
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sc, rays)
    r.exit()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
