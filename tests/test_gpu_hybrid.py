"""GPU: the hybrid execution strategy against the wavefront strategy.

Hybrid runs a step of fresh camera rays (flat, untextured scenes without
ReSTIR) as one traversal-and-shading kernel; every other step, and every
step of other jobs, takes the split kernels.  The reference holds its
strategies to bit-exact agreement (README, execution strategies), and so do
these tests: images, named products and statistics equal the wavefront
strategy's, alone and in a batch that mixes eligible and ineligible jobs.
"""

import sys

import numpy as np
import pytest

from ordinarylight_amd import workloads
from ordinarylight_amd.batch import BatchEngine, EngineSettings, Job, plain_render, run_batch
from ordinarylight_amd.renderers.hip import DeviceContext, DeviceScene, HipPathTracer, HipRendererConfig
from ordinarylight_amd.scene import Material, Transform

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def lights_scene():
    scene, cam = workloads.box_room(2)
    glass = Material(base_color=(0.9, 0.95, 1.0), transmission=1.0, ior=1.5, roughness=0.0)
    v, i = workloads._box((0.4, 0.4, 0.4), None)
    scene.add_mesh(v, i, glass, transform=Transform.translation((0.1, 0.7, 0.3)))
    scene.add_point_light((0.3, 1.6, 0.4), (1.0, 0.9, 0.8), 2.0)
    scene.add_spot_light((-0.5, 1.8, 0.5), (0.3, -1.0, -0.2), (0.8, 0.9, 1.0), 3.0)
    return scene, cam


CASES = {
    "room": (lambda: workloads.box_room(0), 96, 72, 4, 5, {}),
    "terrain": (lambda: workloads.terrain(3, nx=161, nz=81, chunk=40, emissive_quads=12), 128, 72, 2, 8, {}),
    "lights": (lights_scene, 80, 60, 3, 6, {"wavefront_unified_secondary_nee": False,
                                            "wavefront_environment_samples": 2}),
    "chunked": (lambda: workloads.box_room(1), 64, 48, 6, 5, {"wavefront_tile_capacity": 64 * 48}),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_hybrid_equals_wavefront(hip_device, case):
    make, w, h, spp, bounces, extra = CASES[case]
    scene, cam = make()
    log(f"hybrid vs wavefront: {case}")
    out = {}
    for strategy in ("wavefront", "hybrid"):
        r = HipPathTracer(max_bounces=bounces, samples_per_pixel=spp, wavefront_execution_strategy=strategy, **extra)
        img = r.render_frame(scene, cam, w, h)
        prods = r.render_products(scene, cam, w, h, outputs=("color", "depth", "normal", "instance_id"))
        out[strategy] = (img, prods, r.last_stats)
        r.close()
    (a, pa, sa), (b, pb, sb) = out["wavefront"], out["hybrid"]
    assert sb.queue_overflow == 0
    assert np.array_equal(a, b)
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k
    assert (sa.primary_rays, sa.extension_rays, sa.shadow_rays) == (sb.primary_rays, sb.extension_rays,
                                                                     sb.shadow_rays)


def test_hybrid_batches_equal_plain(hip_device):
    """A batch of eligible jobs (their first step runs fused, all jobs in one
    kernel) and a batch that adds a two-level, a textured and a ReSTIR job
    (no step is eligible): every job equals its plain wavefront render."""
    ctx = DeviceContext(hip_device)
    room, room_cam = workloads.box_room(3)
    terrain, terrain_cam = workloads.terrain(5, nx=81, nz=41, chunk=20, emissive_quads=4)
    inst, inst_cam = workloads.instanced_spheres(1, meshes=2, subdivisions=1, instances=20, box=8.0,
                                                 dielectric=0.2, emitters=10, emitter_size=0.3)
    tex, tex_cam = workloads.textured_room(2)
    scenes = [DeviceScene(ctx, room, 2), DeviceScene(ctx, terrain, 2), DeviceScene(ctx, inst, 2, instanced=True),
              DeviceScene(ctx, tex, 2)]
    restir = HipRendererConfig(max_bounces=4, wavefront_restir_di=True, wavefront_restir_candidates=4)
    eligible = [Job(scenes[0], room_cam, 64, 48, 3, 0, None, 5), Job(scenes[1], terrain_cam, 80, 40, 2, 1, None, 6),
                Job(scenes[1], terrain_cam, 24, 24, 1, 5, None, 8), Job(scenes[0], room_cam, 40, 40, 2, 7, None, 3)]
    mixed = eligible + [Job(scenes[2], inst_cam, 48, 48, 2, 2, None, 4), Job(scenes[0], room_cam, 32, 32, 4, 3, restir, 4),
                        Job(scenes[3], tex_cam, 40, 30, 2, 4, None, 3)]
    for name, jobs in (("eligible", eligible), ("mixed", mixed)):
        for fused in (False, True):
            log(f"{name} batch, fused_primary={fused}")
            with BatchEngine(ctx, EngineSettings(path_slots=1 << 18, fused_primary=fused)) as eng:
                _, results = run_batch(eng, jobs, copy_images=True)
                st = eng.stats()
            assert st["queue_overflow"] == 0
            assert st["fused_steps"] == (1 if fused and name == "eligible" else 0)
            for k, (job, (img, _)) in enumerate(zip(jobs, results)):
                alone, _ = plain_render(job)
                assert np.array_equal(img, alone), f"{name} job {k} differs from its render alone (fused {fused})"
    for s in scenes:
        s.close()
    ctx.close()
