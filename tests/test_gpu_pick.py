"""GPU: primary rays and picking.

``olhip_camera_rays`` must produce the render's own primary rays bit for bit
(the oracle generates them independently, wavefront_generate.comp:86-167),
and a pick (targets/vulkan/api.py:936-987) must return the object, triangle,
distance and barycentrics of the pixel's frame-0, sample-0 primary hit: the
primitive equals the oracle's closest hit on the same ray and the primitive
product of a render at that pixel.
"""

import sys

import numpy as np
import pytest

from ordinarylight_amd import cameras, packing, workloads
from ordinarylight_amd.renderers.hip import DeviceContext, HipPathTracer
from ordinarylight_amd.selection import PickOptions, ViewportMapping
from ordinarylight_amd.cameras import OrthographicCamera, PanoramicCamera
from oracle import native as oracle

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx(hip_device):
    c = DeviceContext(hip_device)
    yield c
    c.close()


def _cameras(cam):
    ortho = OrthographicCamera(position=cam.position, target=cam.target, vertical_size=6.0)
    pano = PanoramicCamera(position=cam.position, target=cam.target)
    return {"perspective": cam, "orthographic": ortho, "panoramic": pano}


@pytest.mark.parametrize("kind", ["perspective", "orthographic", "panoramic"])
def test_camera_rays_equal_oracle(ctx, kind):
    _, cam = workloads.box_room(1)
    cam = _cameras(cam)[kind]
    basis, proj = cameras.camera_basis(cam), cameras.camera_projection(cam)
    for frame, sample, tile in ((0, 0, None), (7, 3, (5, 9, 31, 17))):
        o, d = ctx.camera_rays(cam, 64, 48, frame, sample, tile)
        ro, rd = oracle.camera_rays(basis, proj, 64, 48, frame, sample, tile)
        assert np.array_equal(o, ro) and np.array_equal(d, rd), (kind, frame, sample)


def test_pick_matches_primary_hit(hip_device):
    scene, cam = workloads.box_room(2)
    r = HipPathTracer(max_bounces=2, device_index=hip_device)
    W, H = 80, 60
    prods = r.render_products(scene, cam, W, H, outputs=("depth", "object_id"), samples=1)
    orc = oracle.OracleScene(packing.pack_scene(scene))
    rng = np.random.default_rng(4)
    picked = 0
    for _ in range(40):
        x, y = float(rng.uniform(0, W)), float(rng.uniform(0, H))
        res = r.pick(scene, cam, (W, H), (x, y))
        ix, iy = int(x), int(y)
        o, d = oracle.camera_rays(cameras.camera_basis(cam), cameras.camera_projection(cam), W, H, 0, 0,
                                  (ix, iy, 1, 1))
        t, u, v, prim = orc.trace(o[:, :3], d[:, :3], 0.001, 1e30)
        if prim[0] == 0xFFFFFFFF:
            assert res is None and prods["object_id"][iy, ix] == 0xFFFFFFFF
            continue
        picked += 1
        assert res.primitive_index == int(prim[0]) and res.distance == float(t[0])
        assert res.object_id == int(prods["object_id"][iy, ix]) and res.distance == float(prods["depth"][iy, ix])
        start, end = scene.object_triangle_range(res.object_id)
        assert start <= res.primitive_index < end and res.triangle_index == res.primitive_index - start
        assert res.object is scene.get_instance(res.object_id)
        assert abs(sum(res.barycentric) - 1.0) < 1e-6 and min(res.barycentric) >= -1e-6
        assert np.allclose(res.position, o[0, :3] + np.float32(t[0]) * d[0, :3], rtol=0, atol=1e-5)
    assert picked > 20
    # Viewport mapping: a letterboxed viewport maps onto the render pixels.
    m = ViewportMapping((200, 100), render_size=(W, H), content_rect=(20.0, 0.0, 160.0, 100.0))
    assert r.pick(scene, cam, (W, H), (5.0, 50.0), mapping=m) is None
    a = r.pick(scene, cam, (W, H), (100.0, 50.0), mapping=m)
    b = r.pick(scene, cam, (W, H), (40.0, 30.0))
    assert (a is None and b is None) or (a.primitive_index == b.primitive_index)
    with pytest.raises(NotImplementedError):
        r.pick(scene, cam, (W, H), (1.0, 1.0), options=PickOptions(transmissive="through"))
    orc.close()
    r.close()
