"""GPU: the batched engine against single-frame renders.

Every job's image must be bit-identical whether it runs alone
(``olhip_render_gi``) or inside a mixed batch, whatever the batch
composition, admission order, sample chunking or step replay mode: each
path depends only on (pixel, sample, frame) and the float64 pixel sums run
in sample order.  This is the equality the serving design rests on.
"""

import sys
import threading
import time

import numpy as np
import pytest

from ordinarylight_amd import _native as N
from ordinarylight_amd import workloads
from ordinarylight_amd.batch import BatchEngine, EngineSettings, Job, RenderServer, plain_render, run_batch
from ordinarylight_amd.renderers.hip import DeviceContext, DeviceScene, HipRendererConfig
from ordinarylight_amd.scene import Material, Transform

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx(hip_device):
    c = DeviceContext(hip_device)
    yield c
    c.close()


@pytest.fixture(scope="module")
def scenes(ctx):
    room, room_cam = workloads.box_room(1)
    troom, troom_cam = workloads.tessellated_room(2, 4000)
    terr, terr_cam = workloads.terrain(4, nx=121, nz=61, chunk=30, emissive_quads=6)
    lit, lit_cam = workloads.box_room(3)
    glass = Material(base_color=(0.9, 0.95, 1.0), transmission=1.0, ior=1.5, roughness=0.0)
    v, i = workloads._box((0.4, 0.4, 0.4), None)
    lit.add_mesh(v, i, glass, transform=Transform.translation((0.1, 0.7, 0.3)))
    lit.add_point_light((0.3, 1.6, 0.4), (1.0, 0.9, 0.8), 2.0)
    inst, inst_cam = workloads.instanced_spheres(3, meshes=2, subdivisions=2, instances=40, box=10.0,
                                                 emitters=30, emitter_size=0.3)
    out = {
        "inst": (DeviceScene(ctx, inst, 2, instanced=True), inst_cam),
        "room": (DeviceScene(ctx, room, 2), room_cam),
        "troom": (DeviceScene(ctx, troom, 2), troom_cam),
        "terrain": (DeviceScene(ctx, terr, 2), terr_cam),
        "lit": (DeviceScene(ctx, lit, 2), lit_cam),
    }
    yield out
    for ds, _ in out.values():
        ds.close()


def make_jobs(scenes):
    spec = [
        ("room", 64, 48, 3, 4, 0, {}),
        ("terrain", 96, 64, 2, 8, 1, {}),
        ("troom", 40, 40, 5, 6, 2, {}),
        ("lit", 48, 36, 2, 5, 0, {"wavefront_unified_secondary_nee": False, "wavefront_secondary_area_light_samples": 2}),
        ("terrain", 33, 17, 7, 3, 5, {}),
        ("room", 20, 90, 1, 8, 9, {"wavefront_environment_samples": 2}),
        ("lit", 64, 64, 4, 6, 3, {}),
        ("troom", 128, 96, 1, 4, 0, {}),
        ("inst", 72, 40, 3, 8, 4, {}),
    ]
    jobs = []
    for key, w, h, spp, bounces, frame, extra in spec:
        ds, cam = scenes[key]
        cfg = HipRendererConfig(max_bounces=bounces, samples_per_pixel=1, **extra)
        jobs.append(Job(ds, cam, w, h, spp, frame, cfg))
    return jobs


def alone(jobs):
    return [plain_render(j)[0] for j in jobs]


@pytest.mark.parametrize("slots,graph", [(1 << 20, True), (1 << 20, False), (16384, True)])
def test_batch_mix_equals_alone(ctx, scenes, slots, graph):
    """All jobs submitted at once; with 16384 slots the engine must admit
    them over time and split the larger ones into sample chunks."""
    jobs = make_jobs(scenes)
    ref = alone(jobs)
    log(f"batch slots={slots} graph={graph}")
    with BatchEngine(ctx, EngineSettings(path_slots=slots, use_graph=graph)) as eng:
        wall, results = run_batch(eng, jobs, copy_images=True)
        stats = eng.stats()
    log(f"  wall {wall * 1e3:.1f} ms, steps {stats['steps']}, max resident {stats['max_resident_jobs']}")
    for k, ((img, js), r) in enumerate(zip(results, ref)):
        assert img.dtype == np.float32 and img.shape == r.shape
        assert np.array_equal(img, r), f"job {k} differs from its single-frame render"
        assert js["primary_rays"] == jobs[k].width * jobs[k].height * jobs[k].samples
    assert stats["jobs_done"] == len(jobs)
    assert stats["queue_overflow"] == 0
    if slots == 16384:
        assert max(js["chunks"] for _, js in results) > 1
    else:
        assert stats["max_resident_jobs"] == len(jobs)


def test_jobs_join_mid_flight(ctx, scenes):
    """Jobs submitted while others are resident join the running queues."""
    jobs = make_jobs(scenes)
    ref = alone(jobs)
    with BatchEngine(ctx, EngineSettings(path_slots=1 << 20)) as eng:
        first = [eng.submit(j) for j in jobs[:3]]
        eng.step()
        eng.step()
        later = [eng.submit(j) for j in jobs[3:]]
        eng.run()
        stats = eng.stats()
        results = [eng.result(i) for i in first + later]
    assert stats["max_resident_jobs"] >= 4
    assert stats["queue_overflow"] == 0
    for k, ((img, _), r) in enumerate(zip(results, ref)):
        assert np.array_equal(img, r), f"job {k} differs"


def test_server_worker_threads(ctx, scenes):
    """Worker-thread serving: submissions from two client threads with gaps."""
    jobs = make_jobs(scenes)
    ref = alone(jobs)
    server = RenderServer(ctx, EngineSettings(path_slots=1 << 18))
    for key, (ds, _) in scenes.items():
        server.scenes[key] = ds
    ids = [None] * len(jobs)

    def client(indices):
        for k in indices:
            j = jobs[k]
            ids[k] = server.engine.submit(j)
            time.sleep(0.002)

    threads = [threading.Thread(target=client, args=(range(0, len(jobs), 2),)),
               threading.Thread(target=client, args=(range(1, len(jobs), 2),))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    results = [server.wait(i, timeout=120) for i in ids]
    server.engine.close()
    for k, ((img, js), r) in enumerate(zip(results, ref)):
        assert np.array_equal(img, r), f"job {k} differs"
        assert js["done_ms"] >= js["admit_ms"] >= js["submit_ms"]


def test_batch_products_equal_plain(ctx, scenes):
    """Named products of jobs in a batch equal those of single renders."""
    from ordinarylight_amd import _native as N
    jobs = make_jobs(scenes)[:4]
    for j in jobs:
        j.products = N.PRODUCT_PRIMARY | N.PRODUCT_VARIANCE
    with BatchEngine(ctx, EngineSettings(path_slots=1 << 20)) as eng:
        ids = [eng.submit(j) for j in jobs]
        eng.run()
        got = [eng.products(i) for i in ids]
        for i in ids:
            eng.result(i)
    for j, g in zip(jobs, got):
        with BatchEngine(ctx, EngineSettings(path_slots=1 << 20)) as single:
            k = single.submit(j)
            single.run()
            ref = single.products(k)
            single.result(k)
        for name in ("depth", "normal", "primitive", "variance"):
            assert np.array_equal(g[name], ref[name]), name


def test_forced_queue_overflow_fails_the_job_and_the_engine_recovers(ctx, scenes):
    """A queue segment held (from the host) below the no-overflow bound makes
    shading drop continuations: the job fails with the reason instead of
    returning a biased image, and a job admitted afterwards into the same
    slots renders exactly as alone (the failed job's queue counts are gone)."""
    ds, cam = scenes["terrain"]
    cfg = HipRendererConfig(max_bounces=4, samples_per_pixel=1)
    big = Job(ds, cam, 96, 64, 2, 0, cfg)      # 12,288 paths in one segment
    small = Job(ds, cam, 16, 16, 1, 0, cfg)    # 256 paths: at most 256 continuations
    ref = plain_render(small)[0]
    log("forced overflow: append_limit=256")
    with BatchEngine(ctx, EngineSettings(path_slots=1 << 20, append_limit=256)) as eng:
        k = eng.submit(big)
        eng.run()
        assert eng.poll(k) == N.JOB_FAILED
        with pytest.raises(RuntimeError, match="ray queue overflow"):
            eng.result(k)
        assert eng.stats()["queue_overflow"] > 0
        k2 = eng.submit(small)
        eng.run()
        img, _ = eng.result(k2)
    assert np.array_equal(img, ref)
