"""GPU: in-place scene updates (refit and top-level rebuild).

A refit keeps the build's leaf order and recomputes every box from the
moved vertices.  With unmoved vertices it must reproduce the build byte for
byte; with moved vertices, closest hits (with their deterministic tie-break)
and any-hit answers do not depend on the BVH, so traces and GI images must
equal those of a scene built from scratch on the new geometry, bit for bit.
"""

import sys

import numpy as np
import pytest

from ordinarylight_amd import _native as N
from ordinarylight_amd import workloads
from ordinarylight_amd.batch import Job, plain_render
from ordinarylight_amd.renderers.hip import (DeviceContext, DeviceScene, HipPathTracer, HipRendererConfig,
                                             _scene_desc)
from ordinarylight_amd.scene import Transform

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx(hip_device):
    c = DeviceContext(hip_device)
    yield c
    c.close()


def rays(n=20000, seed=3, half=3.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-half, half, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d


def same_trace(a, b, o, d):
    ha, hb = a.trace(o, d, 0.001, 1e30), b.trace(o, d, 0.001, 1e30)
    return all(np.array_equal(x, y) for x, y in zip(ha, hb)), float((ha[3] != 0xFFFFFFFF).mean())


def render(ds, cam, frame=0):
    img, _ = plain_render(Job(ds, cam, 64, 48, 2, frame, HipRendererConfig(max_bounces=5)))
    return img


def test_refit_of_unmoved_vertices_reproduces_the_build(ctx):
    scene, _ = workloads.tessellated_room(1, 5000)
    ds = DeviceScene(ctx, scene, 2)
    nodes0, tri0 = ds.bvh_arrays()
    desc = _scene_desc(ds.packed, 2)
    N.check(ds._lib.olhip_scene_update(ctx.handle, ds.handle, N.ctypes.byref(desc)))
    nodes1, tri1 = ds.bvh_arrays()
    assert np.array_equal(nodes0, nodes1) and np.array_equal(tri0, tri1)
    ds.close()


def test_deformed_mesh_refit_equals_rebuild(ctx):
    scene, cam = workloads.tessellated_room(2, 5000)
    ds = DeviceScene(ctx, scene, 2)
    rng = np.random.default_rng(9)
    resource = scene.mesh_resources[0]
    v = resource.vertices + rng.normal(0.0, 0.01, resource.vertices.shape).astype(np.float32)
    scene.update_mesh_resource(resource, vertices=v)
    assert ds.update(scene) and ds.timings["bvh_refit"]
    fresh = DeviceScene(ctx, scene, 2)
    o, d = rays(half=1.0)
    same, hit = same_trace(ds, fresh, o, d)
    log(f"deformed: hit fraction {hit:.3f}, refit nodes {ds.info.node_count} vs rebuild {fresh.info.node_count}")
    assert same and hit > 0.5
    assert np.array_equal(render(ds, cam), render(fresh, cam))
    # A structural change is refused without touching the resident copy.
    scene.add_mesh(*workloads._box((0.2, 0.2, 0.2), None))
    assert not ds.update(scene)
    ds.close()
    fresh.close()


def test_instance_motion_and_resource_deformation_equal_rebuild(ctx):
    scene, cam = workloads.instanced_spheres(4, meshes=3, subdivisions=2, instances=50, box=10.0, dielectric=0.2,
                                             emitters=20, emitter_size=0.3)
    ds = DeviceScene(ctx, scene, 2, instanced=True)
    rng = np.random.default_rng(2)
    movers = scene.meshes[:20]
    transforms = []
    for inst in movers:
        m = inst.transform.matrix.copy()
        m[:3, 3] += rng.uniform(-1.0, 1.0, 3)
        transforms.append(Transform(m))
    scene.update_instance_transforms(movers, transforms)
    res = scene.mesh_resources[1]
    scene.update_mesh_resource(res, vertices=res.vertices * np.float32(1.1))
    assert ds.update(scene)
    fresh = DeviceScene(ctx, scene, 2, instanced=True)
    o, d = rays(half=5.0)
    same, hit = same_trace(ds, fresh, o, d)
    log(f"instanced update: hit fraction {hit:.3f}")
    assert same and hit > 0.2
    assert np.array_equal(render(ds, cam), render(fresh, cam))
    ds.close()
    fresh.close()


def test_renderer_updates_resident_scene_in_place(hip_device):
    scene, cam = workloads.tessellated_room(3, 3000)
    a = HipPathTracer(max_bounces=4, device_index=hip_device)
    a.render_frame(scene, cam, 48, 32)
    resource = scene.mesh_resources[0]
    scene.update_mesh_resource(resource, vertices=resource.vertices * np.float32(0.98))
    img = a.render_frame(scene, cam, 48, 32)
    assert a.last_timings.get("bvh_refit") is True
    b = HipPathTracer(max_bounces=4, device_index=hip_device)
    assert np.array_equal(img, b.render_frame(scene, cam, 48, 32))
    a.close()
    b.close()


def test_material_and_normal_update_equals_rebuild(ctx):
    # Shading reads each hit's normals and material from per-triangle records
    # built beside the hierarchy: an update that changes only materials and
    # vertex normals must refresh them (the traces are unchanged).
    from ordinarylight_amd.scene import Material
    scene, cam = workloads.tessellated_room(4, 5000)
    ds = DeviceScene(ctx, scene, 2)
    before = render(ds, cam)
    inst = scene.meshes[0]
    scene.update_instance(inst, material=Material(base_color=(0.9, 0.2, 0.1), metallic=1.0, roughness=0.3))
    resource = inst.resource
    rng = np.random.default_rng(5)
    n = resource.normals + rng.normal(0.0, 0.2, resource.normals.shape).astype(np.float32)
    scene.update_mesh_resource(resource, normals=n)
    assert ds.update(scene)
    fresh = DeviceScene(ctx, scene, 2)
    after = render(ds, cam)
    log(f"material/normal update: mean {before[..., :3].mean():.4f} -> {after[..., :3].mean():.4f}")
    assert not np.array_equal(before, after)
    assert np.array_equal(after, render(fresh, cam))
    ds.close()
    fresh.close()
