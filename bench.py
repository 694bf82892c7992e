"""Benchmark: wavefront GI on one MI355X -- plain frames, batched jobs, served jobs.

1. Plain (``value``; one step = one frame): seeded fBm heightfield terrain of
   1,000,000 triangles (1001 x 501 vertices), chunked GGX materials
   (roughness ~ U[0.05,1], metallic ~ Bernoulli(0.3)), 64 emissive triangles
   plus a constant environment of radiance 0.5; 1920 x 1080, 8 samples per
   pixel, 8 bounces, float32.  Mrays/s = (primary + extension + shadow rays
   traced) / wall time of the timed frames (scene build excluded, image kept
   in HBM).  Every scene's BVH is built with the host binned-SAH hierarchy
   (``--builder sah``; ``lbvh`` selects the GPU Morton build); its build time
   is reported beside the rate.
2. Batched (``batched``): 64 jobs, square images with log-uniform edge in
   [128, 1024], eight resident scenes of 10k-2M triangles, spp in
   {1,2,4,8,16}, bounces in {4,8}; one engine run (autotuned capacity and
   launch shape) against the same jobs rendered back-to-back by the plain path.
3. Served (``served``): 256 such jobs (another seed).  Their plain
   back-to-back throughput and the engine's throughput with every job queued
   at once are measured first; then Poisson arrivals at twice the plain
   throughput through the continuous-batching server (jobs/s, p50/p99
   latency; a plain one-at-a-time server on the same arrivals beside it),
   and a sweep of arrival rates at 0.5, 1 and 1.5 times the all-queued
   throughput.
4. Instanced (``c3``): a two-level scene of 10,000 instances of five
   20,480-triangle icospheres (~205M effective triangles), 20 % glass,
   100k emitters sampled with ReSTIR DI (32 candidates, temporal reuse + one
   spatial neighbour); 1920 x 1080, 4 spp, 16 bounces.

Prints one JSON line.  ``--gpus N`` with N > 1 exits: this build measures one GPU.
"""

from __future__ import annotations

import argparse
import dataclasses
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NODE_BYTES, TRI_BYTES, RAY_IN, HIT_OUT, OCC_OUT = 64, 48, 32, 16, 4
# Instance record read when a ray enters an instance: world-to-object rows,
# normal matrix, bounding sphere, bottom-level index and offsets (128 B).
INST_BYTES = 128
# Shading, per closest-hit ray: ray record (origin, direction, path id) and
# hit record in, radiance out; per non-fresh path: throughput, radiance and
# state in (a fresh path's state is rebuilt from its slot); per hit: the
# 64-byte shading record and the 192-byte material; per continuation:
# throughput and state out plus the next ray record; per shadow ray: its
# 48-byte entry and a 32-byte group out (one entry per group without point
# lights); per area-light sample: the 80-byte emitter record.
SHADE_RAY, SHADE_STATE, SHADE_RADIANCE = 36 + 16, 48, 16
SHADE_RECORD, SHADE_MATERIAL, SHADE_CONT = 64, 192, 32 + 36
SHADE_ENTRY, SHADE_GROUP, SHADE_AREA = 48, 32, 80
TRAFFIC_C2 = os.path.join("profiles", "traffic_terrain1M_1080p_8spp_8b.txt")
TRAFFIC_C3 = os.path.join("profiles", "traffic_instanced205M_1080p_4spp_16b.txt")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def host_cpu():
    """lscpu's model, logical CPUs, threads per core and NUMA nodes."""
    import subprocess
    info = {"model": cpu_model(), "cpus": os.cpu_count(), "threads_per_core": None, "numa_nodes": None}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return info
    fields = {}
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            fields[k.strip()] = v.strip()
    info["model"] = fields.get("Model name", info["model"])
    for key, name in (("CPU(s)", "cpus"), ("Thread(s) per core", "threads_per_core"), ("NUMA node(s)", "numa_nodes"),
                      ("Socket(s)", "sockets"), ("Core(s) per socket", "cores_per_socket")):
        try:
            info[name] = int(fields[key])
        except (KeyError, ValueError):
            pass
    return info


def measured_traffic(path, workload, kernel):
    """HBM bytes per dispatch of ``kernel`` from a committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary, when it was measured on this same workload."""
    try:
        with open(os.path.join(ROOT, path)) as f:
            fields = dict((k.strip(), v.strip()) for k, v in
                          (line.split(":", 1) for line in f if line.startswith("traffic ")))
    except OSError:
        return None
    if fields.get("traffic workload") != workload or f"traffic {kernel}" not in fields:
        return None
    return int(fields[f"traffic {kernel}"])


def traversal_bytes(w):
    """Algorithmic bytes of a frame's traversal launches from its work
    counters: node, triangle and instance-record fetches, ray records in,
    hit records or occlusion bits out."""
    closest = w["primary_rays"] + w["extension_rays"]
    return (w["closest_nodes"] * NODE_BYTES + w["closest_tris"] * TRI_BYTES + w["closest_instances"] * INST_BYTES
            + closest * (RAY_IN + HIT_OUT)
            + w["shadow_nodes"] * NODE_BYTES + w["shadow_tris"] * TRI_BYTES + w["shadow_instances"] * INST_BYTES
            + w["shadow_rays"] * (RAY_IN + OCC_OUT))


def shade_bytes(w, instanced=False, face_params=False):
    """Algorithmic bytes of a frame's shading launches (constants above);
    ``face_params``: each hit also reads its triangle's 16-byte base colour
    and roughness."""
    closest = w["primary_rays"] + w["extension_rays"]
    hit = SHADE_RECORD + SHADE_MATERIAL + (INST_BYTES if instanced else 0) + (16 if face_params else 0)
    return (closest * (SHADE_RAY + SHADE_RADIANCE) + w["extension_rays"] * (SHADE_STATE + SHADE_CONT)
            + w["shaded_hits"] * hit + w["shadow_rays"] * (SHADE_ENTRY + SHADE_GROUP)
            + w["area_samples"] * SHADE_AREA)


def roofline_of(kernel, bytes_per_launch, avg_ms, traffic, traffic_source):
    """HBM roofline of one kernel from its algorithmic bytes per launch and
    its average launch time (HIP events on the render stream); ``traffic`` is
    the counter-measured HBM bytes per launch (or None)."""
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel,
           "basis": "algorithmic bytes per launch / average launch time",
           "avg_launch_ms": round(avg_ms, 4), "bytes_per_launch": int(bytes_per_launch),
           "traffic_source": traffic_source if traffic is not None else None}
    if traffic is not None:
        gbs = traffic / (avg_ms * 1e-3) / 1e9
        out["traffic_gbs"] = round(gbs, 1)
        out["traffic_frac"] = round(gbs / HBM_PEAK_GBS, 4)
    return out


def cpu_baseline(scene, camera, width, height, spp, bounces, target_seconds):
    """Time the float32 C oracle on a band of image rows on all host cores."""
    from ordinarylight_amd import cameras, packing
    from oracle import native as oracle
    threads = os.cpu_count() or 1
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    started = time.perf_counter()
    orc = oracle.OracleScene(packing.pack_scene(scene))
    build_s = time.perf_counter() - started
    basis = cameras.camera_basis(camera)
    proj = cameras.camera_projection(camera)
    rows = 4
    elapsed = 0.0
    rays = 0
    row0 = height // 2
    while True:
        r0 = max(0, row0 - rows // 2)
        r1 = min(height, r0 + rows)
        t0 = time.perf_counter()
        orc.render(basis, proj, width, height, spp, 0, max_bounces=bounces, rows=(r0, r1))
        elapsed = time.perf_counter() - t0
        c = orc.last_counts
        rays = int(c[0] + c[1] + c[2])
        if elapsed >= target_seconds * 0.5 or r1 - r0 >= height:
            break
        rows = min(height, max(rows * 2, int(rows * target_seconds / max(elapsed, 1e-3))))
    orc.close()
    host = host_cpu()
    return {
        "value": rays / elapsed / 1e6, "unit": "Mrays/s", "cores": int(os.environ.get("OMP_NUM_THREADS", threads)),
        "kind": "port",
        "sample": f"rows {r0}-{r1} of {width}x{height}, {spp} spp, {bounces} bounces "
                  f"({rays} rays incl. shadow in {elapsed:.1f} s; oracle BVH build {build_s:.1f} s)",
        "host": host,
    }


def c3_cpu_baseline(scene, camera, target_seconds):
    """Time the float32 C oracle's two-level path (instances intersected in
    object space, as the device does) with the C3 settings on all host cores,
    at a reduced 16:9 resolution sized to about ``target_seconds``: ReSTIR
    passes need whole images, so the sample is fewer pixels rather than a
    band of rows."""
    from ordinarylight_amd import cameras, packing
    from oracle import native as oracle
    threads = os.cpu_count() or 1
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    started = time.perf_counter()
    orc = oracle.OracleScene(packing.pack_instanced_scene(scene))
    build_s = time.perf_counter() - started
    basis, proj = cameras.camera_basis(camera), cameras.camera_projection(camera)

    def run(k):
        t0 = time.perf_counter()
        orc.render_restir(basis, proj, 16 * k, 9 * k, 4, 0, candidates=32, spatial_neighbors=1, spatial_radius=4,
                          pairwise=True, history_limit=20, max_bounces=16, environment_samples=1)
        return time.perf_counter() - t0, int(orc.last_counts.sum())

    k = 8
    elapsed, rays = run(k)
    k = int(min(120, max(k, k * math.sqrt(target_seconds / max(elapsed, 1e-3)))))
    elapsed, rays = run(k)
    orc.close()
    return {
        "value": rays / elapsed / 1e6, "unit": "Mrays/s", "cores": int(os.environ.get("OMP_NUM_THREADS", threads)),
        "kind": "port",
        "sample": f"{16 * k}x{9 * k} (C3 scene, camera and settings: 4 ReSTIR passes, 16 bounces; {rays} rays "
                  f"incl. shadow in {elapsed:.1f} s; oracle build {build_s:.1f} s)",
        "host": host_cpu(),
    }


def build_mixed_scenes(context, keys=None, builder="sah"):
    """The eight resident scenes of the batched/served workloads (built once,
    kept resident)."""
    from ordinarylight_amd import workloads
    from ordinarylight_amd.renderers.hip import DeviceScene
    scenes = {}
    for key, _, _ in workloads.MIXED_SCENES:
        if keys is not None and key not in keys:
            continue
        scene, camera = workloads.mixed_scene(key)
        scenes[key] = (DeviceScene(context, scene, 2, builder=builder), camera)
    return scenes


def make_jobs(specs, scenes):
    from ordinarylight_amd.batch import Job
    return [Job(scenes[j["scene"]][0], scenes[j["scene"]][1], j["width"], j["height"], j["samples"],
                0, None, j["max_bounces"]) for j in specs]


def batched_measurement(context, scenes, n_jobs, seed, autotune_on=True):
    """64 mixed jobs: plain back-to-back vs one batched engine run."""
    from ordinarylight_amd import batch, workloads
    jobs = make_jobs(workloads.mixed_jobs(seed, n_jobs), scenes)
    rays = 0
    # Warm-up of both paths (frame-engine growth, graph capture, allocations).
    for j in sorted(jobs, key=lambda j: -j.width * j.height * j.samples)[:2]:
        batch.plain_render(j, copy_image=False)
    t0 = time.perf_counter()
    for j in jobs:
        _, st = batch.plain_render(j, copy_image=False)
        rays += st["primary_rays"] + st["extension_rays"] + st["shadow_rays"]
    plain_s = time.perf_counter() - t0
    if autotune_on:
        settings, timings = batch.autotune(context, jobs)
        tuned = {f"{k.path_slots}/{k.lds_stack}/{k.trace_block}": round(v * 1e3, 2) for k, v in timings.items()}
    else:
        settings, tuned = batch.EngineSettings(), {}
    with batch.BatchEngine(context, settings) as engine:
        batch.run_batch(engine, jobs)  # warm-up
        st0 = engine.stats()
        steps0, dev0 = st0["steps"], st0["device_ms"]
        batched_s, results = batch.run_batch(engine, jobs)
        est = engine.stats()
    batched_rays = sum(st["primary_rays"] + st["extension_rays"] + st["shadow_rays"] for _, st in results)
    # The same batch under the hybrid strategy (its first step, all camera
    # rays, runs as one traversal-and-shading kernel).
    hybrid = dataclasses.replace(settings, fused_primary=True)
    with batch.BatchEngine(context, hybrid) as engine:
        batch.run_batch(engine, jobs)
        hybrid_s, _ = batch.run_batch(engine, jobs)
        fused_steps = engine.stats()["fused_steps"]
    return settings, {
        "jobs": n_jobs,
        "jobs_per_s": round(n_jobs / batched_s, 2),
        "plain_jobs_per_s": round(n_jobs / plain_s, 2),
        "speedup_vs_plain": round(plain_s / batched_s, 3),
        "batched_ms": round(batched_s * 1e3, 2),
        "plain_ms": round(plain_s * 1e3, 2),
        "mrays_per_s": round(batched_rays / batched_s / 1e6, 1),
        "plain_mrays_per_s": round(rays / plain_s / 1e6, 1),
        "max_resident_jobs": int(est["max_resident_jobs"]),
        "engine_steps": int(est["steps"] - steps0),
        "device_busy_frac": round((est["device_ms"] - dev0) / (batched_s * 1e3), 3),
        "hybrid_jobs_per_s": round(n_jobs / hybrid_s, 2), "hybrid_fused_steps": int(fused_steps),
        "settings": {"path_slots": settings.path_slots, "lds_stack": settings.lds_stack,
                     "trace_block": settings.trace_block},
        "autotune_ms": tuned,
    }


def c4_phase(context, scenes, n_jobs, phase):
    """One path of the mixed workload alone (profiling aid): the 64 jobs back
    to back on the plain path, or in one batched engine run with the settings
    the autotuner picks for them on this device (2^27 path slots, 16 LDS stack
    entries, 256-thread traversal blocks); one untimed warm-up pass first."""
    from ordinarylight_amd import batch, workloads
    jobs = make_jobs(workloads.mixed_jobs(0, n_jobs), scenes)
    if phase == "plain":
        for j in jobs:
            batch.plain_render(j, copy_image=False)
        t0 = time.perf_counter()
        for j in jobs:
            batch.plain_render(j, copy_image=False)
        return {"phase": phase, "jobs": n_jobs, "jobs_per_s": round(n_jobs / (time.perf_counter() - t0), 2)}
    settings = batch.EngineSettings(path_slots=1 << 27, lds_stack=16, trace_block=256)
    with batch.BatchEngine(context, settings) as engine:
        batch.run_batch(engine, jobs)
        st0 = engine.stats()
        seconds, _ = batch.run_batch(engine, jobs)
        est = engine.stats()
    return {"phase": phase, "jobs": n_jobs, "jobs_per_s": round(n_jobs / seconds, 2),
            "engine_steps": int(est["steps"] - st0["steps"]),
            "device_busy_frac": round((est["device_ms"] - st0["device_ms"]) / (seconds * 1e3), 3)}


def served_run(context, jobs, settings, rate, seed, check=8, plain_server=False):
    """Poisson arrivals of ``jobs`` at ``rate``/s through the continuous-batching
    server; with ``plain_server`` also a plain one-job-at-a-time server on the
    same arrivals."""
    from ordinarylight_amd import batch
    n_jobs = len(jobs)
    arrivals = batch.poisson_arrivals(rate, n_jobs, seed)
    engine = batch.BatchEngine(context, settings)
    # A warm server: the job list once through the engine first, so step
    # graphs are captured and per-job buffers sit in the engine's pool.
    batch.run_batch(engine, jobs)
    est0 = engine.stats()
    engine.start()
    ids = []
    t0 = time.perf_counter()
    for j, at in zip(jobs, arrivals):
        delay = at - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids.append((engine.submit(j), time.perf_counter()))
    lat, results = [], []
    first_submit, last_done = None, None
    for k, (jid, _) in enumerate(ids):
        state = engine.wait(jid, 600.0)
        if state != 2:
            raise RuntimeError(f"served job {jid} ended in state {state}")
        img, st = engine.result(jid, copy_image=k < check)
        lat.append(st["done_ms"] - st["submit_ms"])
        first_submit = st["submit_ms"] if first_submit is None else min(first_submit, st["submit_ms"])
        last_done = st["done_ms"] if last_done is None else max(last_done, st["done_ms"])
        if k < check:
            results.append(img)
    est = engine.stats()
    engine.close()
    served = batch.latency_percentiles(lat)
    equal = all(np.array_equal(img, batch.plain_render(j)[0]) for img, j in zip(results, jobs[:check]))
    out = {
        "jobs": n_jobs, "rate_per_s": round(rate, 2),
        "jobs_per_s": round(n_jobs / ((last_done - first_submit) * 1e-3), 2),
        "p50_ms": round(served["p50_ms"], 2), "p99_ms": round(served["p99_ms"], 2), "max_ms": round(max(lat), 2),
        "results_equal_plain": bool(equal), "checked_jobs": check,
        "engine_steps": int(est["steps"] - est0["steps"]),
        "device_busy_frac": round((est["device_ms"] - est0["device_ms"]) / (last_done - first_submit), 3),
        "max_resident_jobs": int(est["max_resident_jobs"]),
    }
    if plain_server:
        plain_lat = []
        t0 = time.perf_counter()
        clock = 0.0
        for j, at in zip(jobs, arrivals):
            now = time.perf_counter() - t0
            if now < at:
                time.sleep(at - now)
            batch.plain_render(j, copy_image=False)
            clock = time.perf_counter() - t0
            plain_lat.append((clock - at) * 1e3)
        plain = batch.latency_percentiles(plain_lat)
        out.update({"plain_server_jobs_per_s": round(n_jobs / clock, 2),
                    "plain_server_p50_ms": round(plain["p50_ms"], 2),
                    "plain_server_p99_ms": round(plain["p99_ms"], 2)})
    return out


def served_measurement(context, scenes, settings, n_jobs, seed, load, sweep_loads=(0.5, 1.0, 1.5)):
    """The served workload on one job list: its plain back-to-back capacity
    and the engine's capacity with every job queued at once, then Poisson
    arrivals at ``load`` times the plain capacity (headline, with a plain
    server on the same arrivals) and a sweep of rates around the engine's
    capacity."""
    from ordinarylight_amd import batch, workloads
    jobs = make_jobs(workloads.mixed_jobs(seed, n_jobs), scenes)
    for j in sorted(jobs, key=lambda j: -j.width * j.height * j.samples)[:2]:
        batch.plain_render(j, copy_image=False)
    t0 = time.perf_counter()
    for j in jobs:
        batch.plain_render(j, copy_image=False)
    plain_cap = n_jobs / (time.perf_counter() - t0)
    with batch.BatchEngine(context, settings) as eng:
        batch.run_batch(eng, jobs[:8])
        queued_s, _ = batch.run_batch(eng, jobs)
    queued_cap = n_jobs / queued_s
    out = served_run(context, jobs, settings, load * plain_cap, seed, plain_server=True)
    out.update({
        "load_vs_plain_capacity": load,
        "plain_jobs_per_s": round(plain_cap, 2),
        "all_queued_jobs_per_s": round(queued_cap, 2),
        "speedup_vs_plain": round(out["jobs_per_s"] / plain_cap, 3),
        "all_queued_speedup_vs_plain": round(queued_cap / plain_cap, 3),
    })
    sweep = []
    for s_load in sweep_loads:
        r = served_run(context, jobs, settings, s_load * queued_cap, seed + 1, check=2)
        r["load_vs_all_queued_capacity"] = s_load
        sweep.append(r)
        log(f"served sweep: {r}")
    out["sweep"] = sweep
    return out


def c3_measurement(context, steps=2, builder="lbvh", cpu_seconds=0.0):
    """Instanced two-level scene: 5 jittered icospheres (20,480 triangles each)
    placed 10,000 times (~205M effective triangles), 20 % glass, 100k small
    emitters sampled with ReSTIR DI (32 candidates, temporal reuse and one
    spatial neighbour); 1920x1080, 4 spp, 16 bounces, float32.  The four
    samples run as successive passes, each reusing the previous pass's
    reservoirs."""
    from ordinarylight_amd import workloads
    from ordinarylight_amd.batch import Job, plain_render
    from ordinarylight_amd.renderers.hip import DeviceScene, HipRendererConfig
    t0 = time.perf_counter()
    scene, camera = workloads.instanced_spheres(0)
    gen_s = time.perf_counter() - t0
    ds = DeviceScene(context, scene, 2, instanced=True, builder=builder)
    cfg = HipRendererConfig(max_bounces=16, samples_per_pixel=4, wavefront_restir_di=True,
                            wavefront_restir_candidates=32, wavefront_restir_spatial_reuse=True,
                            wavefront_restir_spatial_neighbors=1)
    job = Job(ds, camera, 1920, 1080, 4, 0, cfg)
    # Work-counting frame (not timed): nodes and triangles per traversal.
    from ordinarylight_amd import _native
    lib = _native.library()
    _native.check(lib.olhip_set_work_counters(context.handle, 1))
    _, work = plain_render(job, copy_image=False)
    _native.check(lib.olhip_set_work_counters(context.handle, 0))
    _native.check(lib.olhip_set_profiling(context.handle, 1))
    plain_render(job, copy_image=False)  # warm-up
    rays, wall = 0, 0.0
    trace_ms = shade_ms = 0.0
    launches = 0
    for _ in range(steps):
        t1 = time.perf_counter()
        _, st = plain_render(job, copy_image=False)
        wall += time.perf_counter() - t1
        rays += st["primary_rays"] + st["extension_rays"] + st["shadow_rays"]
        trace_ms += st["trace_ms"]
        shade_ms += st["shade_ms"]
        launches += st["trace_launches"]
    _native.check(lib.olhip_set_profiling(context.handle, 0))
    workload = "instanced 205M triangles, 1920x1080, 4 spp, 16 bounces, ReSTIR DI"
    roofline = None
    if launches and trace_ms > 0:
        per_frame = launches / steps
        roofline = roofline_of("k_trace2", traversal_bytes(work) / per_frame, trace_ms / launches,
                               measured_traffic(TRAFFIC_C3, workload, "k_trace2"), TRAFFIC_C3)
    closest = max(work["primary_rays"] + work["extension_rays"], 1)
    out = {
        "workload": "5 icospheres x 10,000 instances (204,900,000 triangles), 20% glass, 100k emitters, "
                    "1920x1080, 4 spp, 16 bounces",
        "mrays_per_s": round(rays / wall / 1e6, 1), "ms_per_frame": round(wall * 1e3 / steps, 2),
        "rays_per_frame": int(rays / steps), "scene_generation_s": round(gen_s, 2),
        "scene_build_ms": round(ds.info.build_ms, 2), "effective_triangles": int(ds.triangle_count),
        "nodes_per_closest_ray": round(work["closest_nodes"] / max(work["primary_rays"] + work["extension_rays"], 1), 2),
        "tris_per_closest_ray": round(work["closest_tris"] / max(work["primary_rays"] + work["extension_rays"], 1), 2),
        "nodes_per_shadow_ray": round(work["shadow_nodes"] / max(work["shadow_rays"], 1), 2),
        "tris_per_shadow_ray": round(work["shadow_tris"] / max(work["shadow_rays"], 1), 2),
        "instances_per_closest_ray": round(work["closest_instances"] / closest, 3),
        "instances_per_shadow_ray": round(work["shadow_instances"] / max(work["shadow_rays"], 1), 3),
        "primary_rays": int(work["primary_rays"]), "extension_rays": int(work["extension_rays"]),
        "shadow_rays": int(work["shadow_rays"]),
        "trace_ms_per_frame": round(trace_ms / steps, 3), "shade_ms_per_frame": round(shade_ms / steps, 3),
        "roofline": roofline, "traffic_workload": workload,
        "direct_lighting": "ReSTIR DI at primary hits (32 candidates, temporal + 1 spatial neighbour, "
                           "history limit 20); light-CDF NEE at later bounces",
    }
    ds.close()
    if cpu_seconds > 0:
        out["cpu_baseline"] = c3_cpu_baseline(scene, camera, cpu_seconds)
    return out


def c1_measurement(repeats=5):
    """Box room of 36 triangles (five walls, two boxes rotated about y, one
    emissive quad of radiance 15), 256x256, 4 spp, 4 bounces, seed 0, frame 0:
    the reference's CPU estimator (``CpuReferenceRenderer`` semantics, PCG64
    stream drawn on the host) run on the GPU, best of ``repeats`` frames after
    one warm-up, beside the NumPy restatement of the same estimator timed the
    same way on one host thread."""
    from oracle import e0_numpy
    from ordinarylight_amd import workloads
    from ordinarylight_amd.renderers.hip import HipReferenceRenderer
    scene, camera = workloads.box_room(0)
    W = H = 256
    r = HipReferenceRenderer(samples_per_pixel=4, max_bounces=4, seed=0)
    r.render_frame(scene, camera, W, H)  # warm-up: scene upload, first launches
    best = float("inf")
    for _ in range(repeats):
        t0 = time.perf_counter()
        img = r.render_frame(scene, camera, W, H)
        best = min(best, time.perf_counter() - t0)
    rays = r.last_timings["rays"]
    r.close()
    e0_numpy.render_scene(scene, camera, W, H, samples=4, max_bounces=4, seed=0)
    cpu_best = float("inf")
    for _ in range(repeats):
        t0 = time.perf_counter()
        ref = e0_numpy.render_scene(scene, camera, W, H, samples=4, max_bounces=4, seed=0)
        cpu_best = min(cpu_best, time.perf_counter() - t0)
    a, b = img[..., :3], ref[..., :3]
    close = np.all(np.abs(a - b) <= 1e-6 + 1e-3 * np.maximum(np.abs(a), np.abs(b)), axis=-1)
    return {
        "workload": "box room 36 triangles, 256x256, 4 spp, 4 bounces, seed 0 (reference CPU estimator)",
        "ms_per_frame": round(best * 1e3, 2), "mrays_per_s": round(rays / best / 1e6, 2),
        "rays_per_frame": int(rays),
        "cpu_numpy_ms_per_frame": round(cpu_best * 1e3, 1),
        "cpu_numpy_mrays_per_s": round(rays / cpu_best / 1e6, 3), "cpu_cores": 1,
        "pixels_within_1e-3": round(float(close.mean()), 6),
        "timing": f"best of {repeats} after one warm-up, host PCG64 draws included",
    }


def builder_comparison(scene, camera, W, H, spp, B, leaf, builder, frames=3):
    """The same frame with the scene's BVH built by the other builder: build
    time, nodes per closest-hit ray and Mrays/s over ``frames`` frames."""
    from ordinarylight_amd.renderers.hip import HipPathTracer
    r = HipPathTracer(max_bounces=B, samples_per_pixel=spp, max_samples_per_call=max(64, spp),
                      bvh_leaf_size=leaf, bvh_builder=builder)
    r.replace_scene(scene)
    build_ms = r._device_scene.info.build_ms
    r.reconfigure(work_counters=True)
    r._render(scene, camera, W, H, spp, 0, keep_on_device=True)
    work = r.last_stats
    r.reconfigure(work_counters=False)
    r._render(scene, camera, W, H, spp, 0, keep_on_device=True)
    rays = 0
    t0 = time.perf_counter()
    for _ in range(frames):
        r._render(scene, camera, W, H, spp, 0, keep_on_device=True)
        st = r.last_stats
        rays += st.primary_rays + st.extension_rays + st.shadow_rays
    wall = time.perf_counter() - t0
    r.close()
    closest = work.primary_rays + work.extension_rays
    return {"builder": builder, "mrays_per_s": round(rays / wall / 1e6, 1), "bvh_build_ms": round(build_ms, 2),
            "nodes_per_closest_ray": round(work.closest_nodes / max(closest, 1), 2)}


def strategy_comparison(scene, camera, W, H, spp, B, leaf, builder, strategy, reference_image, frames=3):
    """The same frame under the other execution strategy (``wavefront``:
    every bounce traced and shaded by separate kernels; ``hybrid``: the
    camera-ray bounce traced and shaded by one kernel): Mrays/s and device
    ms per frame over ``frames`` frames, trace/shade split, and whether its
    image equals the headline strategy's bit for bit."""
    from ordinarylight_amd.renderers.hip import HipPathTracer
    r = HipPathTracer(max_bounces=B, samples_per_pixel=spp, max_samples_per_call=max(64, spp),
                      bvh_leaf_size=leaf, bvh_builder=builder, wavefront_execution_strategy=strategy, profiling=True)
    r.replace_scene(scene)
    image = r.render_frame(scene, camera, W, H)
    rays, dev, trace, shade = 0, 0.0, 0.0, 0.0
    t0 = time.perf_counter()
    for _ in range(frames):
        r._render(scene, camera, W, H, spp, 0, keep_on_device=True)
        st = r.last_stats
        rays += st.primary_rays + st.extension_rays + st.shadow_rays
        dev, trace, shade = dev + st.total_ms, trace + st.trace_ms, shade + st.shade_ms
    wall = time.perf_counter() - t0
    r.close()
    return {"strategy": strategy, "mrays_per_s": round(rays / wall / 1e6, 1),
            "device_ms_per_frame": round(dev / frames, 3), "trace_ms_per_frame": round(trace / frames, 3),
            "shade_ms_per_frame": round(shade / frames, 3),
            "image_equal": bool(reference_image is not None and np.array_equal(image, reference_image))}


def main(argv=None):
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=5)
    parser.add_argument("--warmup", type=int, default=1)
    parser.add_argument("--width", type=int, default=1920)
    parser.add_argument("--height", type=int, default=1080)
    parser.add_argument("--spp", type=int, default=8)
    parser.add_argument("--bounces", type=int, default=8)
    parser.add_argument("--cpu-seconds", type=float, default=12.0)
    parser.add_argument("--no-cpu", action="store_true")
    parser.add_argument("--no-profile", action="store_true")
    parser.add_argument("--leaf-size", type=int, default=2)
    parser.add_argument("--builder", choices=("lbvh", "sah"), default="sah",
                        help="BVH hierarchy of every scene (build time is reported, not timed)")
    parser.add_argument("--lds-stack", type=int, default=16)
    parser.add_argument("--block", type=int, default=128)
    parser.add_argument("--batch-jobs", type=int, default=64, help="batched workload size (0: skip)")
    parser.add_argument("--serve-jobs", type=int, default=256, help="served workload size (0: skip)")
    parser.add_argument("--serve-load", type=float, default=2.0,
                        help="Poisson arrival rate as a multiple of the plain path's back-to-back jobs/s")
    parser.add_argument("--no-autotune", action="store_true")
    parser.add_argument("--no-c3", action="store_true", help="skip the instanced two-level workload")
    parser.add_argument("--no-c1", action="store_true", help="skip the reference-estimator box room")
    parser.add_argument("--no-builder-comparison", action="store_true",
                        help="skip the plain frame with the other BVH builder")
    parser.add_argument("--strategy", choices=("wavefront", "hybrid"), default="wavefront",
                        help="execution strategy of the plain frame (the other one is measured beside it)")
    parser.add_argument("--no-strategy-comparison", action="store_true",
                        help="skip the plain frame under the other execution strategy")
    parser.add_argument("--c3-only", action="store_true",
                        help="run only the instanced workload and print its record (profiling aid)")
    parser.add_argument("--c4-only", choices=("plain", "batched"), default=None,
                        help="run only the mixed 64-job workload on one path and print its record (profiling aid)")
    args = parser.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(the path shards by image tiles or whole jobs, one GPU per share)", file=sys.stderr)
        return 2

    from ordinarylight_amd import workloads
    from ordinarylight_amd.renderers.hip import HipPathTracer

    if args.c4_only:
        from ordinarylight_amd.renderers.hip import DeviceContext
        context = DeviceContext(0)
        scenes = build_mixed_scenes(context, builder=args.builder)
        print(json.dumps({"c4": c4_phase(context, scenes, args.batch_jobs or 64, args.c4_only)}), flush=True)
        for ds, _ in scenes.values():
            ds.close()
        context.close()
        return 0

    if args.c3_only:
        from ordinarylight_amd.renderers.hip import DeviceContext
        context = DeviceContext(0)
        print(json.dumps({"c3": c3_measurement(context, steps=args.steps, builder=args.builder,
                                               cpu_seconds=0.0 if args.no_cpu else args.cpu_seconds)}), flush=True)
        context.close()
        return 0

    t0 = time.perf_counter()
    scene, camera = workloads.terrain(0)
    log(f"scene: {scene.triangle_count} triangles, generated in {time.perf_counter() - t0:.1f} s")
    W, H, spp, B = args.width, args.height, args.spp, args.bounces
    renderer = HipPathTracer(max_bounces=B, samples_per_pixel=spp, max_samples_per_call=max(64, spp),
                             bvh_leaf_size=args.leaf_size, bvh_builder=args.builder, traversal_lds_stack=args.lds_stack,
                             traversal_block=args.block, wavefront_execution_strategy=args.strategy)
    # Scene upload + BVH build (reported separately, outside the timed region).
    renderer.replace_scene(scene)
    ds = renderer._device_scene
    log(f"bvh: {ds.info.node_count} nodes, depth {ds.info.bvh_depth}, build {ds.info.build_ms:.1f} ms")

    # Work-counting pass (not timed): nodes/triangles per traversal.
    renderer.reconfigure(work_counters=True)
    renderer._render(scene, camera, W, H, spp, 0, keep_on_device=True)
    work = renderer.last_stats
    renderer.reconfigure(work_counters=False, profiling=not args.no_profile)

    for _ in range(args.warmup):
        renderer._render(scene, camera, W, H, spp, 0, keep_on_device=True)
    rays = 0
    dev_ms = 0.0
    trace_ms = shade_ms = 0.0
    trace_launches = 0
    t_start = time.perf_counter()
    for step in range(args.steps):
        renderer._render(scene, camera, W, H, spp, 0, keep_on_device=True)
        st = renderer.last_stats
        rays += st.primary_rays + st.extension_rays + st.shadow_rays
        dev_ms += st.total_ms
        trace_ms += st.trace_ms
        shade_ms += st.shade_ms
        trace_launches += st.trace_launches
    wall_s = time.perf_counter() - t_start
    ms_per_step = wall_s * 1000.0 / args.steps
    value = rays / (wall_s * 1e6)

    # Achievable streaming bandwidth on this device (1 GiB read + 1 GiB write).
    import ctypes
    from ordinarylight_amd import _native
    ms = ctypes.c_double(0)
    _native.check(_native.library().olhip_bandwidth_probe(renderer._context.handle, 1 << 30, 10, ctypes.byref(ms)))
    copy_gbs = 2 * (1 << 30) / (ms.value * 1e-3) / 1e9

    # PCIe-inclusive: one frame with the image copied back to the host.
    t1 = time.perf_counter()
    headline_image = renderer.render_frame(scene, camera, W, H)
    pcie_ms = (time.perf_counter() - t1) * 1000.0

    # The traversal launch of a step traces its closest hits and the previous
    # step's shadow rays: its algorithmic bytes are both rays' node and
    # triangle fetches, ray records in, hit records or occlusion bits out.
    closest_rays = work.primary_rays + work.extension_rays
    w = work.as_dict()
    roofline = shade_roofline = None
    workload = (f"terrain 1M triangles, per-triangle GGX, {W}x{H}, {spp} spp, {B} bounces, "
                f"leaf {args.leaf_size}, {args.builder}")
    if trace_ms > 0 and trace_launches:
        launches_per_frame = trace_launches / args.steps
        roofline = roofline_of("k_trace", traversal_bytes(w) / launches_per_frame, trace_ms / trace_launches,
                               measured_traffic(TRAFFIC_C2, workload, "k_trace"), TRAFFIC_C2)
        faces = ds.packed.face_params is not None
        shade_roofline = roofline_of("k_shade<false>", shade_bytes(w, face_params=faces) / launches_per_frame,
                                     shade_ms / trace_launches,
                                     measured_traffic(TRAFFIC_C2, workload, "k_shade<false>"), TRAFFIC_C2)
    line = {
        "metric": "Mrays/s at 1080p 8-bounce; batched jobs/s + p50/p99 latency vs plain GPU",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded fBm terrain)",
        "config": {"workload": "terrain 1M triangles, per-triangle GGX materials, 1920x1080, 8 spp, 8 bounces, "
                               "64 emissive tris + env 0.5",
                   "bvh_builder": args.builder, "execution_strategy": args.strategy,
                   "width": W, "height": H, "spp": spp, "max_bounces": B, "triangles": int(ds.info.triangle_count)},
        "roofline": roofline,
        "extra": {
            "rays_per_frame": int(rays / args.steps),
            "device_ms_per_frame": round(dev_ms / args.steps, 3),
            "trace_ms_per_frame": round(trace_ms / args.steps, 3),
            "shade_ms_per_frame": round(shade_ms / args.steps, 3),
            "copy_probe_gbs": round(copy_gbs, 1),
            "pcie_inclusive_frame_ms": round(pcie_ms, 3),
            "bvh_build_ms": round(ds.info.build_ms, 2),
            "bvh_nodes": int(ds.info.node_count),
            "nodes_per_closest_ray": round(work.closest_nodes / max(closest_rays, 1), 2),
            "tris_per_closest_ray": round(work.closest_tris / max(closest_rays, 1), 2),
            "nodes_per_shadow_ray": round(work.shadow_nodes / max(work.shadow_rays, 1), 2),
            "tris_per_shadow_ray": round(work.shadow_tris / max(work.shadow_rays, 1), 2),
            "shadow_rays_per_frame": int(work.shadow_rays),
            "shaded_hits_per_frame": int(work.shaded_hits),
            "shade_roofline": shade_roofline,
            "traffic_workload": workload,
        },
    }
    renderer.close()
    if not args.no_strategy_comparison:
        other = "hybrid" if args.strategy == "wavefront" else "wavefront"
        line["extra"]["other_strategy"] = strategy_comparison(scene, camera, W, H, spp, B, args.leaf_size,
                                                              args.builder, other, headline_image)
        log(f"other strategy: {line['extra']['other_strategy']}")
    if not args.no_builder_comparison:
        other = "lbvh" if args.builder == "sah" else "sah"
        line["extra"]["other_builder"] = builder_comparison(scene, camera, W, H, spp, B, args.leaf_size, other)
        log(f"other builder: {line['extra']['other_builder']}")
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(scene, camera, W, H, spp, B, args.cpu_seconds)
    if not args.no_c1:
        line["c1"] = c1_measurement()
        log(f"c1: {line['c1']}")

    if not args.no_c3:
        from ordinarylight_amd.renderers.hip import DeviceContext
        context = DeviceContext(0)
        line["c3"] = c3_measurement(context, builder=args.builder, cpu_seconds=0.0 if args.no_cpu else args.cpu_seconds)
        log(f"c3: {line['c3']}")
        context.close()

    if args.batch_jobs > 0 or args.serve_jobs > 0:
        from ordinarylight_amd.renderers.hip import DeviceContext
        context = DeviceContext(0)
        t0 = time.perf_counter()
        scenes = build_mixed_scenes(context, builder=args.builder)
        log(f"mixed scenes: {len(scenes)} resident in {time.perf_counter() - t0:.1f} s")
        settings = None
        if args.batch_jobs > 0:
            settings, line["batched"] = batched_measurement(context, scenes, args.batch_jobs, 0,
                                                            not args.no_autotune)
            log(f"batched: {line['batched']}")
        if args.serve_jobs > 0:
            from ordinarylight_amd.batch import EngineSettings
            # One job list: arrivals at --serve-load times its plain
            # back-to-back capacity (headline), then a load sweep around the
            # engine's capacity on the same list.
            line["served"] = served_measurement(context, scenes, settings or EngineSettings(), args.serve_jobs, 1,
                                                args.serve_load)
            log(f"served: {line['served']}")
        for ds, _ in scenes.values():
            ds.close()
        context.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
