"""Benchmark: incoherent closest-hit throughput of the extend stage.

Workload: the 256-instance TLAS scene over 4 BLAS (value-noise heightfield
256x256 quads, icosphere subdiv 6, 200K-triangle soup, 12-quad box room) plus
4 emissive quads, 413,016 unique / ~26M instanced triangles, camera
(0,5,-30), 1920x1080.  Rays: 2^21 incoherent rays (origin uniform in the
scene bounds, direction uniform on the sphere, seed 3).  One step = one
closest-hit pass over all rays, inputs resident in HBM.

Prints one JSON line (metric, value, unit, steps, config ...) with a roofline object for the extend
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


PROFILE = os.path.join(ROOT, "profiles", "extend_tlas256_incoherent.md")


PROFILE_KEYS = ("traffic_bytes_per_launch", "valu_busy", "lane_utilisation", "l2_hit", "waves_per_simd",
                "salu_per_valu")


def measured_profile(cfg: dict):
    """Counters of the extend launch on this workload from the committed
    rocprofv3 summary: HBM bytes per launch (FETCH_SIZE and WRITE_SIZE
    scaled by a plain copy's counts), VALU busy, lane utilisation, L2 hit
    rate, resident waves per SIMD and SALU per VALU instruction; None when
    the summary was taken with another traversal configuration."""
    try:
        text = open(PROFILE).read()
    except OSError:
        return None
    want = " ".join(f"{k}={v}" for k, v in cfg.items())
    got, vals = None, {}
    for line in text.splitlines():
        if line.startswith("traversal_config:"):
            got = line.split(":", 1)[1].strip()
        key = line.split(":", 1)[0]
        if key in PROFILE_KEYS:
            vals[key] = float(line.split(":", 1)[1])
    if got != want or "traffic_bytes_per_launch" not in vals:
        return None
    vals["source"] = os.path.relpath(PROFILE, ROOT)
    return vals


def parse_traversal(text: str):
    """'block=64,lds_stack=16' -> {'block': 64, 'lds_stack': 16}; '' -> None."""
    if not text:
        return None
    return {k.strip(): int(v) for k, v in (kv.split("=") for kv in text.split(","))}


def host_threads() -> int:
    """Every CPU thread this process may use: the affinity mask, capped by
    OMP_NUM_THREADS when the environment grants fewer (a shared host's CPU
    share per GPU)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(omp)) if omp.isdigit() and int(omp) > 0 else n)


def host_description() -> dict:
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    import subprocess
    try:
        cc = subprocess.run(["gcc", "--version"], capture_output=True, text=True, timeout=10).stdout.splitlines()[0]
    except (OSError, IndexError, subprocess.SubprocessError):
        cc = "unknown"
    return {"cpu_model": model, "host_cpus": os.cpu_count(), "compiler": cc,
            "flags": "-O3 -std=c11 -ffp-contract=off -fno-fast-math"}


def _sample_rate(orc, rays, threads, seconds_target):
    """Oracle closest-hit passes over a prefix of `rays` sized to about
    `seconds_target` of work: (Mrays/s, rays in the sample, passes, seconds)."""
    n = 4096
    while True:
        t0 = time.perf_counter()
        orc.trace_closest(rays[:n], nthreads=threads)
        dt = time.perf_counter() - t0
        if dt > 1.0 or n >= len(rays):
            break
        n = min(len(rays), n * 4)
    passes = max(1, int(round(seconds_target / max(dt, 1e-6))))
    n_final = n
    if n < len(rays):
        n_final = int(min(len(rays), n * seconds_target / max(dt, 1e-6)))
        passes = 1
    sample = rays[:n_final]
    t0 = time.perf_counter()
    for _ in range(passes):
        orc.trace_closest(sample, nthreads=threads)
    dt = time.perf_counter() - t0
    return passes * n_final / dt / 1e6, n_final, passes, dt


def cpu_baseline(scene, rays, seconds_target=12.0):
    """The oracle's closest-hit traversal on the host: on every CPU thread
    the process is given (the value) and on one thread, whole passes over
    the same incoherent rays until about `seconds_target` of work each."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_api as O
    orc = O.OracleScene(scene)
    threads = host_threads()
    v, n_final, passes, dt = _sample_rate(orc, rays, threads, seconds_target)
    v1, n1, p1, d1 = _sample_rate(orc, rays, 1, seconds_target / 3)
    return {"value": round(v, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "one_thread_mrays": round(v1, 4), **host_description(),
            "sample": f"{passes} pass(es) over the first {n_final} of the {len(rays)} incoherent rays, oracle traversal "
                      f"(C, -O3, no FP contraction), {threads} threads, {dt:.1f} s; one thread: {p1} pass(es) over "
                      f"the first {n1}, {d1:.1f} s"}


def time_cpu(orc, rays, threads, seconds_target):
    """Oracle closest-hit passes over `rays` until ~seconds_target of work."""
    t0 = time.perf_counter()
    orc.trace_closest(rays, nthreads=threads)
    dt = time.perf_counter() - t0
    passes = max(1, int(seconds_target / max(dt, 1e-6)))
    t0 = time.perf_counter()
    for _ in range(passes):
        orc.trace_closest(rays, nthreads=threads)
    dt = time.perf_counter() - t0
    return round(passes * len(rays) / dt / 1e6, 4), passes, round(dt, 1)


def bench_bounce1(r, sc, stream, steps: int, warmup: int):
    """The 256-instance scene's other ray set: first-bounce rays (cosine-
    distributed about the hit normals) of the 1920x1080 camera rays, closest
    hit on the renderer `r` already holding the scene."""
    import torch
    from wavefrontpathtracing_amd import scenes
    from wavefrontpathtracing_amd import _lib as L
    cam = scenes.primary_rays(sc)
    b1 = scenes.bounce_rays(sc, cam, r.trace_closest(cam))
    d_b1 = torch.from_numpy(b1.view(np.uint8)).to("cuda:0")
    d_b1h = torch.empty(len(b1) * 20, dtype=torch.uint8, device="cuda:0")
    for _ in range(warmup):
        r.trace_closest_device(d_b1.data_ptr(), len(b1), d_b1h.data_ptr())
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(steps):
        r.trace_closest_device(d_b1.data_ptr(), len(b1), d_b1h.data_ptr())
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    hits = d_b1h.cpu().numpy().view(L.HIT)
    return {"rays": len(b1), "mrays": round(len(b1) / (ms * 1e-3) / 1e6, 2), "ms_per_step": round(ms, 4),
            "hit_fraction": round(float((hits["instance_idx"] != L.MISS).mean()), 4),
            "source": "cosine-distributed directions about the hit normals of the 1920x1080 camera rays"}


def bench_soup(n_tris: int, steps: int, cpu: bool):
    """Random-triangle soup (centroids in [-1,1]^3, offsets +-U(0.005,0.02)),
    16-bin SAH, camera (0,0,-3) -> origin, 60 deg, 512x512: closest hit of
    the 262,144 camera rays and of 2^21 incoherent rays (uniform origin in the
    soup bounds, uniform sphere directions)."""
    import torch
    from wavefrontpathtracing_amd import Renderer, scenes
    sc = scenes.soup_scene(n_tris)
    r = Renderer(sc.width, sc.height)
    r.set_build_options(interval_count=16)
    scenes.load(r, sc)
    r.build_tlas()
    bmin, bmax = r.scene_bounds()
    sets = {"primary": scenes.primary_rays(sc), "incoherent": scenes.sphere_rays(3, bmin, bmax, 1 << 21)}
    dev = torch.device("cuda:0")
    stream = torch.cuda.ExternalStream(r.stream)
    out = {"workload": f"{n_tris:,}-triangle soup, 16-bin SAH, 512x512 camera rays + 2^21 incoherent rays"}
    for name, rays in sets.items():
        d_rays = torch.from_numpy(rays.view(np.uint8)).to(dev)
        d_hits = torch.empty(len(rays) * 20, dtype=torch.uint8, device=dev)
        torch.cuda.synchronize()
        for _ in range(2):
            r.trace_closest_device(d_rays.data_ptr(), len(rays), d_hits.data_ptr())
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            r.trace_closest_device(d_rays.data_ptr(), len(rays), d_hits.data_ptr())
        e1.record(stream)
        torch.cuda.synchronize()
        out[f"{name}_mrays"] = round(len(rays) * steps / (e0.elapsed_time(e1) * 1e-3) / 1e6, 2)
    r.exit()
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_api as O
        orc = O.OracleScene(sc, bins=16)
        threads = host_threads()
        prim = sets["primary"]
        v1, p1, d1 = time_cpu(orc, prim[: len(prim) // 8], 1, 4.0)
        vn, pn, dn = time_cpu(orc, prim, threads, 4.0)
        out["cpu_primary"] = {"1_thread_mrays": v1, f"{threads}_threads_mrays": vn, "kind": "port",
                              "sample": f"oracle traversal: 1 thread {p1} pass(es) over 1/8 of the camera rays "
                                        f"({d1} s); {threads} threads {pn} pass(es) over all ({dn} s)"}
    return out


def bench_build(cpu: bool):
    """BLAS build on the GPU (wf_bvh_builder_build_device, scratch kept across
    builds) over random soups: 65,536 triangles at 16 intervals (the CPU
    reference case), 1M at 16 and 2M at 8 (the largest meshes of the batch).
    Each device build is checked equal, node for node and index for index, to
    the host builder's; the host builder (threads) and the oracle's restated
    reference builder (one thread) are timed beside it."""
    import torch
    from wavefrontpathtracing_amd import renderer as R, scenes
    b = R.DeviceBvhBuilder()
    out = {"workload": "random soups (centroids in [-1,1]^3, offsets +-U(0.005,0.02)), SAH binned BLAS; "
                       "GPU time = wall clock of the whole build call, positions already in HBM"}
    stream = torch.cuda.current_stream().cuda_stream
    for n, bins, key in ((65536, 16, "soup_65k_16"), (1 << 20, 16, "soup_1m_16"), (2_000_000, 8, "soup_2m_8")):
        v, i = scenes.random_soup(n % 1000 + 1, n)
        pos9 = np.ascontiguousarray(v["position"][i].reshape(-1, 9), dtype=np.float32)
        d_pos = torch.from_numpy(pos9).to("cuda")
        d_nodes = torch.empty((2 * n, 8), dtype=torch.int32, device="cuda")
        d_idx = torch.empty(n, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        args = (d_pos.data_ptr(), n, d_nodes.data_ptr(), 2 * n, d_idx.data_ptr(), bins, False, stream)
        b.build(*args)
        reps = 10 if n <= 65536 else 5
        t0 = time.perf_counter()
        for _ in range(reps):
            count = b.build(*args)
        gpu_s = (time.perf_counter() - t0) / reps
        levels, subtrees, max_leaf = b.last_build()
        t0 = time.perf_counter()
        blob = R.bvh_build(pos9, bins)
        host_s = time.perf_counter() - t0
        nodes_h, _, idx_h = R.parse_bvh_blob(blob)
        assert count == len(nodes_h)
        assert d_nodes[:count].cpu().numpy().tobytes() == nodes_h.tobytes(), "device BLAS nodes differ from the host build"
        assert d_idx.cpu().numpy().tobytes() == idx_h.tobytes(), "device BLAS indices differ from the host build"
        e = {"tris": n, "intervals": bins, "gpu_ms": round(gpu_s * 1e3, 3), "gpu_mtris_per_s": round(n / gpu_s / 1e6, 1),
             "host_builder_ms": round(host_s * 1e3, 1),
             "host_builder": "the library's host builder, large subtrees of the recursion on std::async threads",
             "nodes": count, "level_passes": levels, "wave_subtrees": subtrees, "max_leaf": max_leaf,
             "equal_to_host_build": True}
        if cpu and n <= (1 << 20):
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            import oracle_api as O
            reps_c = 20 if n <= 65536 else 1
            t0 = time.perf_counter()
            for _ in range(reps_c):
                O.blas_blob(pos9, bins)
            c_s = (time.perf_counter() - t0) / reps_c
            e["cpu_oracle_ms"] = round(c_s * 1e3, 1)
            e["cpu_oracle"] = {"kind": "port", "cores": 1,
                               "sample": f"{reps_c} build(s) of the same soup by the oracle's restated builder"}
            e["speedup_vs_cpu_oracle"] = round(c_s / gpu_s, 1)
        out[key] = e
        del d_pos, d_nodes, d_idx
    b.close()
    # TLAS: the agglomerative build over random instances (the 256-instance
    # scene's transform distribution, spread so the density stays constant)
    from wavefrontpathtracing_amd import _lib as L
    bmin, bmax = np.array([-1, -0.5, -1], np.float32), np.array([1, 0.5, 1], np.float32)
    for m in (256, 4096, 32767):
        rng = scenes.Rng(m)
        spread = 10.0 * max(1.0, (m / 256) ** (1 / 3))
        inst = np.array([R.make_bvh_instance(rng.transform(0.5, 2.0, spread), bmin, bmax, k % 4) for k in range(m)],
                        dtype=L.BVH_INSTANCE)
        R.tlas_build_device(inst)
        reps = 5 if m <= 4096 else 1
        t0 = time.perf_counter()
        for _ in range(reps):
            dblob = R.tlas_build_device(inst)
        gpu_s = (time.perf_counter() - t0) / reps
        e = {"instances": m, "gpu_ms": round(gpu_s * 1e3, 2)}
        if m <= 4096 or cpu:
            t0 = time.perf_counter()
            hblob = R.tlas_build(inst)
            host_s = time.perf_counter() - t0
            assert dblob == hblob, "device TLAS differs from the host build"
            e.update({"host_ms": round(host_s * 1e3, 2), "host_kind": "the library's host build (one thread)",
                      "equal_to_host_build": True, "speedup_vs_host": round(host_s / gpu_s, 2)})
        out[f"tlas_{m}"] = e
    return out


def bench_frame(frames: int, union: bool = True):
    """One wavefront frame of the 256-instance scene at 1920x1080, 1 spp,
    depth 5, Russian roulette from bounce 3, next-event estimation to the 4
    emissive quads, black environment: stage times from the library's own
    events (direct launches; shadow rays traced in the next bounce's extend
    launch, so ms_extend includes them), then the same frame as a captured
    graph, with one launch per stage, and as one megakernel; with `union`,
    both modes again with the union-area TLAS."""
    from wavefrontpathtracing_amd import Renderer, scenes
    sc = scenes.tlas_scene()
    r = Renderer(sc.width, sc.height)
    scenes.load(r, sc)
    r.update_settings(max_bounces=5, spp=1, next_event_estimation=1, hdr_env_strength=0.0, frame_seed=0,
                      rr_start_bounce=3, rr_survival=0.75, accumulate=1, use_hip_graph=0)
    r.render()
    keys = ("ms_clear", "ms_generate", "ms_extend", "ms_shade", "ms_connect", "ms_post", "ms_total")
    acc = {k: 0.0 for k in keys}
    rays = 0
    for _ in range(frames):
        r.render()
        st = r.frame_stats()
        for k in keys:
            acc[k] += getattr(st, k)
        rays += st.rays_extended + st.shadow_rays
    r.update_settings(use_hip_graph=1)
    r.render()
    gt = 0.0
    for _ in range(frames):
        r.render()
        gt += r.frame_stats().ms_total
    # one launch per stage (shadow rays in their own connect launch)
    r.update_settings(use_hip_graph=0, separate_connect=1)
    r.render()
    sep = {k: 0.0 for k in ("ms_extend", "ms_connect", "ms_total")}
    for _ in range(frames):
        r.render()
        st = r.frame_stats()
        for k in sep:
            sep[k] += getattr(st, k)
    # the same frame as one megakernel launch (use_wavefront_pathtracing = 0)
    r.update_settings(separate_connect=0, use_wavefront_pathtracing=0)
    r.render()
    mt = 0.0
    for _ in range(frames):
        r.render()
        mt += r.frame_stats().ms_total
    # both modes with the union-metric TLAS; one frame at a fixed seed with
    # each TLAS checks that the image is the same
    r.update_settings(frame_seed=7)
    r.render()
    ref_img = r.read_energy()
    um = None
    if union:
        r.set_tlas_metric("union")
        r.build_tlas()
        r.render()
        um = {"image_identical": bool(np.array_equal(r.read_energy(), ref_img))}
        r.update_settings(frame_seed=0)
        for wavefront in (1, 0):
            r.update_settings(use_wavefront_pathtracing=wavefront)
            r.render()
            t = 0.0
            for _ in range(frames):
                r.render()
                t += r.frame_stats().ms_total
            um["wavefront_ms_total" if wavefront else "megakernel_ms_total"] = round(t / frames, 3)
    r.exit()
    out = {k: round(v / frames, 3) for k, v in acc.items()}
    out.update({"workload": "256-instance scene, 1920x1080, 1 spp, depth 5, RR from 3, NEE, black environment",
                "rays_per_frame": rays // frames, "mrays_per_s": round(rays / (acc["ms_total"] * 1e-3) / 1e6, 2),
                "graph_ms_total": round(gt / frames, 3), "megakernel_ms_total": round(mt / frames, 3),
                "separate_connect": {k: round(v / frames, 3) for k, v in sep.items()},
                "union_tlas": um})
    return out


def bench_inverse(steps: int):
    """Batched inverse rendering: 64 orbit views of the 256-instance scene at
    512x512, 8 spp, depth 5, NEE, a 1024x1024 albedo texture and the 4 light
    strengths as unknowns; each step = forward of all views in one batch +
    path-replay adjoint + Adam update."""
    from wavefrontpathtracing_amd import inverse as INV
    t0 = time.perf_counter()
    hist = INV.run(steps=steps + 1, views=64, res=512, spp=8, max_bounces=5, tex_size=1024, lr=1e-2, seed=7, log=None)
    total = time.perf_counter() - t0
    timed = hist[1:]
    s = float(np.mean([h["s"] for h in timed]))
    return {"workload": "64 views x 512x512, 8 spp, depth 5, NEE, 256-instance scene (union-area TLAS), 1024x1024 "
                        "texture + 4 emission scalars, forward + adjoint + Adam per step",
            "ms_per_step": round(s * 1e3, 1), "steps_per_s": round(1.0 / s, 3),
            "forward_rays_per_step": timed[-1]["forward_rays"],
            "loss": [round(h["loss"], 6) for h in hist], "texel_mae": [round(h["texel_mae"], 5) for h in hist],
            "steps_timed": steps, "wall_s_with_setup": round(total, 1)}


def batch_candidates(headline):
    """The configurations the batch tuner times (the same list for the batch
    and for the per-scene loop): the default, the four fastest on the
    headline rays, and the fastest there with a scratch (private) stack,
    with the 4-wide layout, and with and without the popped-interior skip
    (2-wide, LDS stack); then the default and the two fastest there at
    refill thresholds of 24 and 32 idle lanes (launches of short rays claim
    queue entries often, and fewer, larger claims can pay)."""
    from wavefrontpathtracing_amd import tuning
    picks = [dict(tuning.DEFAULT)]
    # frames do not order their queues by cost: that axis stays at its default here
    ranked = [{**c, "cost_order": tuning.DEFAULT["cost_order"]} for c, _ in headline] if headline \
        else list(tuning.candidates())
    picks += ranked[:4]
    for pred in (lambda c: c["lds_stack"] == 0, lambda c: c["bvh_layout"] == 1,
                 lambda c: c["bvh_layout"] == 0 and c["lds_stack"] > 0 and c["skip_popped"] == 1,
                 lambda c: c["bvh_layout"] == 0 and c["lds_stack"] > 0 and c["skip_popped"] == 0):
        picks += [c for c in ranked if pred(c)][:1]
    for c in [dict(tuning.DEFAULT)] + ranked[:2]:
        picks += [dict(c, refill_min=m) for m in (24, 32)]
    out, seen = [], set()
    for c in picks:
        key = tuning.label(c)
        if key not in seen:
            seen.add(key)
            out.append(c)
    return out


def bench_batched(frames: int, tune: bool, headline=None, fixed=None, batch_only: bool = False):
    """64 uneven scenes (soup of 10K..2M triangles in a lit box room, 256^2 /
    512^2 / 1024^2 pixels), 4 spp, depth 4, next-event estimation: the whole
    batch as one frame of one launch per stage (captured hipGraph) against
    the plain loop of one frame per scene.  Both produce the same images
    (the batch path numbering equals the per-scene one).  `fixed` (a
    traversal configuration) replaces the tuner for both.  `batch_only`
    stops after the batched frames (for profiling them)."""
    from wavefrontpathtracing_amd import Renderer, scenes, tuning
    t0 = time.perf_counter()
    sl = scenes.batch_scenes()
    r = Renderer(256, 256)
    scenes.load_batch(r, sl)
    setup_s = time.perf_counter() - t0
    r.update_settings(max_bounces=4, spp=4, next_event_estimation=1, hdr_env_strength=0.0, frame_seed=9,
                      rr_start_bounce=3, rr_survival=0.75, accumulate=0, use_hip_graph=1)
    n = len(sl)

    def batched_frame():
        t = time.perf_counter()
        r.render_scenes(0, n)
        dt = time.perf_counter() - t
        st = r.frame_stats()
        return dt, st.rays_extended + st.shadow_rays

    def plain_pass():
        t = time.perf_counter()
        for k in range(n):
            r.render_scenes(k, 1)
        return time.perf_counter() - t

    if fixed:
        tuning.set_extend_config(r, **fixed)
        tune = False
    tuned = None
    if tune:
        t_tune = time.perf_counter()
        best, table = tuning.autotune(r, lambda: batched_frame()[0], configs=batch_candidates(headline), repeats=1)
        tuned = {"winner": best, "seconds": round(time.perf_counter() - t_tune, 1),
                 "table_ms_per_frame": {tuning.label(c): round(t * 1e3, 1) for c, t in table},
                 "labels": "block/lds_stack/skip_popped/bvh_layout/interior_visits/refill_min/cost_order"}
    batched_frame()
    bt, brays = 0.0, 0
    for _ in range(frames):
        dt, nr = batched_frame()
        bt += dt
        brays += nr
    if batch_only:
        r.exit()
        return {"batched_mrays": round(brays / bt / 1e6, 2), "batched_ms_per_frame": round(bt / frames * 1e3, 2),
                "rays_per_frame": brays // frames, "frames": frames, "fixed_traversal": fixed, "batch_only": True}
    batch_imgs = [r.read_scene_energy(k, sc.width, sc.height) for k, sc in enumerate(sl)]
    # the same images, one scene at a time (no graph), with the traversal
    # configuration tuned for that loop
    r.update_settings(use_hip_graph=0)
    tuned_plain = None
    if tune:
        t_tune = time.perf_counter()
        best_p, table_p = tuning.autotune(r, plain_pass, configs=batch_candidates(headline), repeats=1)
        tuned_plain = {"winner": best_p, "seconds": round(time.perf_counter() - t_tune, 1),
                       "table_ms_per_pass": {tuning.label(c): round(t * 1e3, 1) for c, t in table_p}}
    for k in range(n):
        r.render_scenes(k, 1)
    pt, prays = 0.0, 0
    for _ in range(frames):
        for k in range(n):
            t = time.perf_counter()
            r.render_scenes(k, 1)
            pt += time.perf_counter() - t
            st = r.frame_stats()
            prays += st.rays_extended + st.shadow_rays
    differ = [k for k, sc in enumerate(sl) if not np.array_equal(r.read_scene_energy(k, sc.width, sc.height), batch_imgs[k])]
    del batch_imgs
    if differ:
        raise RuntimeError(f"batched and per-scene images differ for scenes {differ[:8]}")
    # the reference's frame structure: one scene, one sample per pixel per
    # frame, frames accumulated until the scene has 4 samples, one launch per
    # stage (shadow rays in their own launch)
    r.update_settings(spp=1, accumulate=1, frame_seed=0, separate_connect=1)
    rt, rrays = 0.0, 0
    for _ in range(frames):
        for k in range(n):
            for _s in range(4):
                t = time.perf_counter()
                r.render_scenes(k, 1)
                rt += time.perf_counter() - t
                st = r.frame_stats()
                rrays += st.rays_extended + st.shadow_rays
    # the batch as one megakernel launch per frame
    r.update_settings(spp=4, accumulate=0, frame_seed=9, use_wavefront_pathtracing=0, separate_connect=0)
    batched_frame()
    mt, mrays = 0.0, 0
    for _ in range(frames):
        dt, nr = batched_frame()
        mt += dt
        mrays += nr
    r.exit()
    b_mr, p_mr, r_mr = brays / bt / 1e6, prays / pt / 1e6, rrays / rt / 1e6
    return {"workload": f"{n} scenes, {sum(len(i) // 3 for s in sl for _, i in s.meshes[:1]):,} soup tris, "
                        f"{sum(s.width * s.height for s in sl):,} pixels, 4 spp, depth 4, NEE",
            "batched_mrays": round(b_mr, 2), "plain_mrays": round(p_mr, 2), "ratio": round(b_mr / p_mr, 3),
            "reference_loop_mrays": round(r_mr, 2), "ratio_vs_reference_loop": round(b_mr / r_mr, 3),
            "megakernel_batched_mrays": round(mrays / mt / 1e6, 2),
            "reference_loop_ms_per_batch": round(rt / frames * 1e3, 2),
            "rays_per_frame": brays // frames, "batched_ms_per_frame": round(bt / frames * 1e3, 2),
            "plain_ms_per_frame": round(pt / frames * 1e3, 2), "setup_s": round(setup_s, 1), "frames": frames,
            "images_identical": True,
            "autotuned": tuned, "autotuned_plain_loop": tuned_plain, "fixed_traversal": fixed}


def bench_adjoint(frames: int):
    """Differentiable room (512^2, 16 spp, depth 4, NEE, 512^2 value-noise
    albedo texture, two subdivision-6 icospheres): forward frame and the
    path-replay adjoint of an L2 loss against a target rendered with another
    albedo texture.  adjoint spp/s = spp / adjoint seconds per frame."""
    import torch
    from wavefrontpathtracing_amd import Renderer, scenes
    sc = scenes.differentiable_scene()
    tgt = scenes.differentiable_scene(tex_seed=77).textures[0][2]
    r = Renderer(sc.width, sc.height)
    _, tex_h = scenes.load(r, sc)
    spp = 16
    r.update_settings(max_bounces=4, spp=spp, next_event_estimation=1, hdr_env_strength=0.0, frame_seed=3,
                      accumulate=0, use_hip_graph=1)
    dev = torch.device("cuda:0")
    n_pix = sc.width * sc.height
    tw, th = sc.textures[0][0], sc.textures[0][1]
    t_src = np.ones((th * tw, 4), np.float32)
    t_src[:, :3] = tgt.reshape(-1, 3)
    r.update_texture(tex_h[0], t_src)
    r.render()
    target = torch.empty((n_pix, 4), dtype=torch.float32, device=dev)
    r.read_energy_device(target.data_ptr())
    t0 = np.ones((th * tw, 4), np.float32)
    t0[:, :3] = sc.textures[0][2].reshape(-1, 3)
    r.update_texture(tex_h[0], t0)
    img = torch.empty_like(target)
    g_tex = torch.empty((th * tw, 4), dtype=torch.float32, device=dev)
    g_em = torch.empty(len(sc.instances), dtype=torch.float32, device=dev)
    fwd, adj = [], []
    for k in range(frames + 1):
        t = time.perf_counter()
        r.render()
        tf = time.perf_counter() - t
        r.read_energy_device(img.data_ptr())
        dl = (2.0 * (img - target)).contiguous()
        torch.cuda.synchronize()
        t = time.perf_counter()
        r.render_adjoint_device(dl.data_ptr(), tex_h[0], g_tex.data_ptr(), g_em.data_ptr())
        ta = time.perf_counter() - t
        if k:
            fwd.append(tf)
            adj.append(ta)
    r.exit()
    tf, ta = float(np.mean(fwd)), float(np.mean(adj))
    return {"workload": f"differentiable room {sc.width}x{sc.height}, {spp} spp, depth 4, NEE, "
                        f"{tw}x{th} albedo texture + emission",
            "adjoint_spp_per_s": round(spp / ta, 2), "forward_spp_per_s": round(spp / tf, 2),
            "adjoint_ms": round(ta * 1e3, 2), "forward_ms": round(tf * 1e3, 2), "cost_ratio": round(ta / tf, 3),
            "grad_norm": float(torch.linalg.vector_norm(g_tex[:, :3]).item()), "frames": frames}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rays", type=int, default=1 << 21)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-soup", action="store_true")
    ap.add_argument("--no-frame", action="store_true")
    ap.add_argument("--no-inverse", action="store_true")
    ap.add_argument("--no-build", action="store_true")
    ap.add_argument("--no-union", action="store_true",
                    help="skip the opt-in union-area TLAS measurements (headline rays and frame)")
    ap.add_argument("--no-bounce1", action="store_true",
                    help="skip the first-bounce ray set (profiled runs: the headline kernel then runs only on the headline rays)")
    ap.add_argument("--build-only", action="store_true",
                    help="only the BLAS/TLAS build section (for profiling the builders); prints its JSON")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner: run the default traversal configuration (or --traversal)")
    ap.add_argument("--traversal", default="",
                    help="fixed traversal configuration, e.g. block=64,lds_stack=16,interior_visits=4 (implies --no-tune)")
    ap.add_argument("--batch-traversal", default="",
                    help="fixed traversal configuration of the batched section alone (its tuner then does not run)")
    ap.add_argument("--batch-frames-only", action="store_true",
                    help="batched section: only the batched frames, no per-scene, reference-loop or megakernel "
                         "comparison (for profiling the batched frame)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy 5 times (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": src.numel() * 4, "launches": 5}))
        return
    if args.build_only:
        print(json.dumps({"bvh_build": bench_build(cpu=not args.no_cpu)}), flush=True)
        return
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

    autotuned = None
    if args.traversal:
        tuning.set_extend_config(r, **parse_traversal(args.traversal))
    elif not args.no_tune:
        # the autotuner over every workgroup size x stack placement x
        # popped-interior skip x node layout, at 2 and 4 interior visits per
        # pass, on this workload; the winner is the configuration timed below
        ev_a, ev_b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

        def timed():
            ev_a.record(stream)
            step()
            ev_b.record(stream)
            ev_b.synchronize()
            return ev_a.elapsed_time(ev_b) * 1e-3

        t_tune = time.perf_counter()
        cfgs = [dict(c, interior_visits=v) for c in tuning.candidates() for v in (2, 4)]
        _, table = tuning.autotune(r, timed, configs=cfgs, repeats=3)
        # the refill threshold for the fastest four, and the callers' order
        # (no cost order) for the default and the fastest two; the default
        # stays unless beaten by 1 %
        cfgs2 = [dict(tuning.DEFAULT)] + [dict(c, refill_min=m) for c, _ in table[:4] for m in (16, 24, 32)]
        cfgs2 += [dict(c, cost_order=2) for c in [dict(tuning.DEFAULT)] + [c for c, _ in table[:2]]]
        best, table2 = tuning.autotune(r, timed, configs=cfgs2, repeats=3)
        seen = {tuning.label(c) for c, _ in table2}
        table = sorted(table2 + [(c, t) for c, t in table if tuning.label(c) not in seen], key=lambda x: x[1])
        autotuned = {"workload": "the headline rays", "candidates": len(table), "seconds": round(time.perf_counter() - t_tune, 1),
                     "winner": best, "table_mrays": {tuning.label(c): round(len(rays) / t / 1e6, 1) for c, t in table},
                     "labels": "block/lds_stack/skip_popped/bvh_layout/interior_visits/refill_min/cost_order"}

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
    # the same rays through the host-array entry point (PCIe copies included)
    r.trace_closest(rays)
    t_host = time.perf_counter()
    for _ in range(3):
        host_hits = r.trace_closest(rays)
    host_mrays = 3 * len(rays) / (time.perf_counter() - t_host) / 1e6
    assert host_hits.tobytes() == hits.tobytes()
    _, counters = r.trace_closest_counters(rays)
    b_ray = bytes_per_ray(counters)
    achieved_gbs = mrays * 1e6 * b_ray / 1e9
    cfg = tuning.get_extend_config(r)
    prof = measured_profile(cfg)
    traffic = None if prof is None else prof["traffic_bytes_per_launch"]

    # the same rays with the union-metric TLAS (opt-in): same hits, fewer node visits
    union = None
    if not args.no_union:
        r.set_tlas_metric("union")
        r.build_tlas()
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        ev0.record(stream)
        for _ in range(args.steps):
            step()
        ev1.record(stream)
        torch.cuda.synchronize()
        u_ms = ev0.elapsed_time(ev1) / args.steps
        u_hits = d_hits.cpu().numpy().view(L.HIT)
        _, u_counters = r.trace_closest_counters(rays)
        union = {"mrays": round(len(rays) / (u_ms * 1e-3) / 1e6, 2), "ms_per_step": round(u_ms, 4),
                 "hits_identical": u_hits.tobytes() == hits.tobytes(),
                 "mean_counters": {k: round(float(v), 3) for k, v in
                                   zip(("tlas_nodes", "blas_nodes", "instances", "tri_tests"), u_counters.mean(axis=0))}}
        r.set_tlas_metric("reference")
        r.build_tlas()

    # the scene's other ray set: first-bounce rays from the hits of the 1920x1080 camera rays
    bounce1 = None
    if not args.no_bounce1:
        bounce1 = bench_bounce1(r, sc, stream, args.steps, args.warmup)

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
            "host_arrays_mrays": round(host_mrays, 2),
            "wall_s_timed": round(t_wall, 4),
            "union_tlas": union,
            "bounce1": bounce1,
            "autotuned": autotuned,
        },
        "roofline": {
            # priced against HBM, the only published peak that applies (no
            # MFMA work); the counters say what limits the kernel instead
            "bound": "hbm",
            "achieved": round(achieved_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved_gbs / HBM_PEAK_GBS, 4),
            "traffic": None if traffic is None else round(traffic),
            "traffic_unit": "B per launch",
            "bytes_per_ray": round(b_ray, 1),
            "achieved_is": "algorithmic bytes (reference-algorithm traversal counters x record sizes) / kernel time",
            "mean_counters": {k: round(float(v), 3) for k, v in
                              zip(("tlas_nodes", "blas_nodes", "instances", "tri_tests"), counters.mean(axis=0))},
            "measured": None if prof is None else {
                "hbm_gbs": round(traffic / (ms_step * 1e-3) / 1e9, 1),
                "hbm_frac": round(traffic / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                **{k: prof[k] for k in PROFILE_KEYS[1:] if k in prof}, "source": prof["source"]},
            "limiter": "issue+divergence",
            "limiter_evidence": None if prof is None else (
                f"measured HBM traffic {traffic / (ms_step * 1e-3) / 1e9:,.0f} GB/s "
                f"({traffic / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS:.3f} of peak; node fetches hit L2 at "
                f"{prof.get('l2_hit', float('nan')):.2f} and the Infinity Cache); VALU busy "
                f"{prof.get('valu_busy', float('nan')):.2f}; lane utilisation {prof.get('lane_utilisation', float('nan')):.2f} "
                f"({64 * prof.get('lane_utilisation', float('nan')):.1f} of 64 lanes active per VALU instruction: SIMT "
                f"divergence and the tail of the longest rays)"),
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sc, rays)
    r.exit()
    if not args.no_frame:
        out["frame"] = bench_frame(frames=5, union=not args.no_union)
    if not args.no_soup:
        out["soup_65k"] = bench_soup(65536, steps=10, cpu=not args.no_cpu)
        out["soup_1m"] = bench_soup(1 << 20, steps=10, cpu=False)
    if not args.no_build:
        out["bvh_build"] = bench_build(cpu=not args.no_cpu)
    if not args.no_batch:
        out["batched"] = bench_batched(frames=2, tune=not args.no_tune and not args.traversal,
                                       headline=None if autotuned is None else table,
                                       fixed=parse_traversal(args.batch_traversal),
                                       batch_only=args.batch_frames_only)
    if not args.no_adjoint:
        out["adjoint"] = bench_adjoint(frames=3)
    if not args.no_inverse:
        out["inverse"] = bench_inverse(steps=3)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
