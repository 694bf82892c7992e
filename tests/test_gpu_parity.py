"""GPU parity: libolhip.so against the oracles on the same inputs.

* BVH build: byte-identical to the CPU LBVH oracle (integer work).
* Closest hit: identical (t, u, v, triangle) to the oracle's independent BVH
  traversal (lexicographic (t, index) minimum).
* GI estimator: per-pixel radiance against the float32 C oracle fed the same
  hash sample streams; tolerance 1e-3 relative (+1e-6 absolute) on at least
  99.9 % of pixels, mismatches confined to path divergences at edges.
* Reference estimator: radiance against the NumPy restatement of
  CpuReferenceRenderer on the same PCG64 stream, 1e-3 relative; primary-hit
  triangle IDs equal on >= 99.99 % of pixels.
"""

import sys

import numpy as np
import pytest

from ordinarylight_amd import cameras, packing, workloads
from ordinarylight_amd.renderers.hip import (
    DeviceContext, DeviceScene, HipPathTracer, HipReferenceRenderer,
)
from oracle import e0_numpy
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


def soup(n, seed):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-10, 10, (n, 3)).astype(np.float32)
    verts = (centers[:, None, :] + rng.normal(0, 0.6, (n, 3, 3)).astype(np.float32)).reshape(-1, 3)
    return verts, np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)


def scene_from_arrays(verts, idx):
    from ordinarylight_amd.scene import Scene
    s = Scene()
    s.add_mesh(verts, idx)
    return s


@pytest.mark.parametrize("case", ["room", "soup", "terrain", "one", "two"])
@pytest.mark.parametrize("leaf", [1, 4])
def test_bvh_bitexact(ctx, case, leaf):
    log(f"bvh case={case} leaf={leaf}")
    if case == "room":
        scene, _ = workloads.box_room(0)
    elif case == "soup":
        scene = scene_from_arrays(*soup(20000, 3))
    elif case == "terrain":
        scene, _ = workloads.terrain(1, nx=121, nz=61, chunk=30, emissive_quads=4)
    elif case == "one":
        scene = scene_from_arrays(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32),
                                  np.array([[0, 1, 2]], np.uint32))
    else:
        scene = scene_from_arrays(*soup(2, 9))
    ds = DeviceScene(ctx, scene, leaf_size=leaf)
    nodes, tri = ds.bvh_arrays()
    ref = oracle.lbvh_build(ds.packed.positions, ds.packed.indices, leaf)
    assert ds.info.bvh_depth == ref["depth"]
    assert nodes.shape == ref["nodes"].shape
    assert np.array_equal(nodes, ref["nodes"])
    assert np.array_equal(tri.view(np.uint32), ref["tri"].view(np.uint32))
    ds.close()


def random_rays(lo, hi, n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d


def test_closest_hit_matches_oracle(ctx):
    scene, cam = workloads.terrain(2, nx=201, nz=101, chunk=50, emissive_quads=8)
    ds = DeviceScene(ctx, scene)
    orc = oracle.OracleScene(ds.packed)
    lo, hi = ds.packed.positions.min(0), ds.packed.positions.max(0)
    o, d = random_rays(lo - 5, hi + 5, 200000, 4)
    # Also rays aimed exactly at vertices and edge midpoints (ties).
    tri = ds.packed.positions[ds.packed.indices]
    rng = np.random.default_rng(5)
    pick = rng.integers(0, len(tri), 20000)
    targets = np.concatenate((tri[pick, 0], 0.5 * (tri[pick, 0] + tri[pick, 1])))
    o2 = np.repeat(np.asarray(cam.position, np.float32)[None], len(targets), 0)
    d2 = targets - o2
    d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
    o = np.concatenate((o, o2)).astype(np.float32)
    d = np.concatenate((d, d2)).astype(np.float32)
    g = ds.trace(o, d, 0.001, 1e30)
    c = orc.trace(o, d, 0.001, 1e30)
    hit_frac = float(np.mean(g[3] != 0xFFFFFFFF))
    log(f"closest-hit rays={len(o)} hit fraction={hit_frac:.3f}")
    assert hit_frac > 0.2
    same_id = g[3] == c[3]
    assert same_id.mean() >= 0.9999, f"triangle ids differ on {np.sum(~same_id)} rays"
    both = same_id & (g[3] != 0xFFFFFFFF)
    assert np.array_equal(g[0][both].view(np.uint32), c[0][both].view(np.uint32))
    assert np.array_equal(g[1][both].view(np.uint32), c[1][both].view(np.uint32))
    assert np.array_equal(g[2][both].view(np.uint32), c[2][both].view(np.uint32))
    orc.close()
    ds.close()


def rel_close(a, b, rel=1e-3, abs_=1e-6):
    return np.all(np.abs(a - b) <= abs_ + rel * np.maximum(np.abs(a), np.abs(b)), axis=-1)


def gi_compare(scene, cam, w, h, spp, bounces, frame=0, **cfg):
    r = HipPathTracer(max_bounces=bounces, samples_per_pixel=spp, max_samples_per_call=4096, **cfg)
    img = r.render_frame(scene, cam, w, h, frame_index=frame)
    stats = r.last_stats
    orc = oracle.OracleScene(packing.pack_scene(scene))
    ref = orc.render(cameras.camera_basis(cam), cameras.camera_projection(cam), w, h, spp, frame,
                     max_bounces=bounces,
                     secondary_area_samples=r.config.wavefront_secondary_area_light_samples or r.config.area_light_samples,
                     environment_samples=r.config.wavefront_environment_samples,
                     roulette_start=r.config.wavefront_russian_roulette_start if r.config.wavefront_russian_roulette else 0,
                     roulette_min_survival=r.config.wavefront_russian_roulette_min_survival,
                     nee_probability=r.config.wavefront_secondary_nee_probability,
                     unified=r.config.wavefront_unified_secondary_nee)
    r.close()
    orc.close()
    assert stats.queue_overflow == 0
    return img, ref, stats


@pytest.mark.parametrize("frame", [0, 7])
def test_gi_room_matches_oracle(hip_device, frame):
    scene, cam = workloads.box_room(0)
    log(f"gi room frame={frame}")
    img, ref, stats = gi_compare(scene, cam, 96, 72, 4, 4, frame=frame)
    assert img.dtype == np.float32 and img.shape == (72, 96, 4)
    assert np.all(img[..., 3] == 1.0)
    ok = rel_close(img[..., :3], ref[..., :3])
    exact = np.all(img == ref, axis=-1)
    log(f"  within tol {ok.mean():.6f} exact {exact.mean():.6f} rays {stats.primary_rays}+{stats.extension_rays}+{stats.shadow_rays}")
    assert ok.mean() >= 0.999
    assert stats.primary_rays == 96 * 72 * 4


def test_gi_terrain_matches_oracle(hip_device):
    scene, cam = workloads.terrain(3, nx=161, nz=81, chunk=40, emissive_quads=12)
    log("gi terrain")
    img, ref, stats = gi_compare(scene, cam, 128, 72, 2, 8)
    ok = rel_close(img[..., :3], ref[..., :3])
    log(f"  within tol {ok.mean():.6f} exact {np.all(img == ref, axis=-1).mean():.6f}")
    assert np.isfinite(img).all()
    assert ok.mean() >= 0.999


def test_gi_lights_transmission_matches_oracle(hip_device):
    """Point/spot/directional lights, a dielectric box and non-unified NEE."""
    from ordinarylight_amd.scene import Material
    scene, cam = workloads.box_room(2)
    glass = Material(base_color=(0.9, 0.95, 1.0), transmission=1.0, ior=1.5, roughness=0.0)
    v, i = workloads._box((0.4, 0.4, 0.4), None)
    from ordinarylight_amd.scene import Transform
    scene.add_mesh(v, i, glass, transform=Transform.translation((0.1, 0.7, 0.3)))
    scene.add_point_light((0.3, 1.6, 0.4), (1.0, 0.9, 0.8), 2.0)
    scene.add_spot_light((-0.5, 1.8, 0.5), (0.3, -1.0, -0.2), (0.8, 0.9, 1.0), 3.0)
    scene.add_directional_light((0.2, -1.0, 0.1), (1, 1, 1), 0.5)
    for unified in (True, False):
        log(f"gi lights unified={unified}")
        img, ref, _ = gi_compare(scene, cam, 80, 60, 3, 6, wavefront_unified_secondary_nee=unified,
                                 wavefront_secondary_area_light_samples=2,
                                 wavefront_environment_samples=2)
        ok = rel_close(img[..., :3], ref[..., :3])
        log(f"  within tol {ok.mean():.6f}")
        assert ok.mean() >= 0.999


def test_gi_many_emitters_matches_oracle(hip_device):
    """3,000 emissive triangles of unequal power: the device narrows the
    emitter bisection with its selection guide, the oracle bisects the whole
    CDF, and both must pick the same emitters."""
    from ordinarylight_amd.scene import Material
    scene, cam = workloads.box_room(4)
    rng = np.random.default_rng(11)
    quad = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)  # facing down
    for _ in range(1500):
        x, z = rng.uniform(-0.95, 0.9, 2)
        s = rng.uniform(0.01, 0.05)
        v = np.array([[x, 1.999, z], [x + s, 1.999, z], [x + s, 1.999, z + s], [x, 1.999, z + s]], np.float32)
        e = float(rng.uniform(0.5, 30.0))
        scene.add_mesh(v, quad, Material(base_color=(0.5, 0.5, 0.5), emission=(e, 0.9 * e, 0.8 * e)))
    log("gi many emitters")
    img, ref, stats = gi_compare(scene, cam, 80, 60, 2, 4)
    ok = rel_close(img[..., :3], ref[..., :3])
    log(f"  within tol {ok.mean():.6f} shadow rays {stats.shadow_rays}")
    assert stats.shadow_rays > 0
    assert ok.mean() >= 0.999


def test_gi_tile_capacity_invariance(hip_device):
    """Chunked residency gives the identical image (tests/gates/tile_quality.py)."""
    scene, cam = workloads.box_room(1)
    a = HipPathTracer(max_bounces=5, samples_per_pixel=6)
    b = HipPathTracer(max_bounces=5, samples_per_pixel=6, wavefront_tile_capacity=64 * 48)
    ia = a.render_frame(scene, cam, 64, 48)
    ib = b.render_frame(scene, cam, 64, 48)
    assert b.last_stats.chunks == 6
    assert np.array_equal(ia, ib)
    a.close()
    b.close()


@pytest.mark.parametrize("case", ["room", "terrain"])
def test_gi_converged_matches_fp64_oracle(hip_device, case):
    """1024 spp on the GPU (float32 paths, float64 pixel sums) against the
    float64 oracle of the same estimator and sample streams: relative mean
    absolute error of the converged image <= 1 %."""
    if case == "room":
        scene, cam = workloads.box_room(4)
        w, h, bounces = 24, 18, 5
    else:
        scene, cam = workloads.terrain(5, nx=101, nz=51, chunk=25, emissive_quads=4)
        w, h, bounces = 24, 16, 8
    spp = 1024
    log(f"converged {case} {w}x{h} spp={spp} bounces={bounces}")
    r = HipPathTracer(max_bounces=bounces, samples_per_pixel=spp, max_samples_per_call=spp)
    img = r.render_frame(scene, cam, w, h)
    r.close()
    orc = oracle.OracleScene(packing.pack_scene(scene), "f64")
    ref = orc.render(cameras.camera_basis(cam), cameras.camera_projection(cam), w, h, spp, 0, max_bounces=bounces)
    orc.close()
    err = float(np.abs(img[..., :3].astype(np.float64) - ref[..., :3]).mean() / np.abs(ref[..., :3]).mean())
    log(f"  relative mean error {err:.2e}")
    assert np.isfinite(img).all()
    assert err <= 0.01


@pytest.mark.parametrize("leaf", [1, 2, 4])
def test_gi_launch_shape_invariance(hip_device, leaf):
    """LDS stack split and block size leave the image bit-identical."""
    scene, cam = workloads.terrain(3, nx=161, nz=81, chunk=40, emissive_quads=6)
    shapes = [(16, 128), (8, 128), (16, 64), (16, 256), (32, 128)]
    ref = None
    for lds, block in shapes:
        log(f"launch shape leaf={leaf} lds={lds} block={block}")
        r = HipPathTracer(max_bounces=6, samples_per_pixel=3, bvh_leaf_size=leaf, traversal_lds_stack=lds,
                          traversal_block=block)
        img = r.render_frame(scene, cam, 96, 64)
        r.close()
        if ref is None:
            ref = img
        else:
            assert np.array_equal(img, ref), (lds, block)


def test_e0_room_matches_numpy_oracle(hip_device):
    scene, cam = workloads.box_room(0)
    log("e0 room")
    r = HipReferenceRenderer(samples_per_pixel=4, max_bounces=4, seed=0)
    img = r.render_frame(scene, cam, 64, 64, frame_index=0, capture_primary_hits=True)
    trace = []
    ref = e0_numpy.render_scene(scene, cam, 64, 64, samples=4, max_bounces=4, seed=0, frame_index=0, trace=trace)
    assert img.dtype == np.float32 and img.shape == (64, 64, 4)
    assert np.all(img[..., 3] == 0.0)
    ok = rel_close(img[..., :3], ref[..., :3])
    log(f"  within tol {ok.mean():.6f} exact {np.all(img == ref, axis=-1).mean():.6f}")
    assert ok.mean() >= 0.9999
    primary = trace[0][3]
    assert np.mean(r.last_primary_hits == primary) >= 0.9999
    r.close()


@pytest.mark.parametrize("kind", ["ortho", "pano", "env"])
def test_e0_projections_match_numpy_oracle(hip_device, kind):
    from ordinarylight_amd.lights import EnvironmentLight
    scene, cam = workloads.box_room(1)
    if kind == "ortho":
        cam = cameras.OrthographicCamera((0.0, 1.0, 3.0), (0.0, 1.0, 0.0), vertical_size=2.2)
    elif kind == "pano":
        cam = cameras.PanoramicCamera((0.0, 1.0, 0.0), (0.0, 1.0, -1.0))
    else:
        scene.set_environment(EnvironmentLight(color=(0.9, 0.8, 0.6), intensity=0.7))
    log(f"e0 {kind}")
    r = HipReferenceRenderer(samples_per_pixel=2, max_bounces=3, seed=5)
    img = r.render_frame(scene, cam, 48, 40, frame_index=3)
    ref = e0_numpy.render_scene(scene, cam, 48, 40, samples=2, max_bounces=3, seed=5, frame_index=3)
    ok = rel_close(img[..., :3], ref[..., :3])
    log(f"  within tol {ok.mean():.6f}")
    assert ok.mean() >= 0.999
    r.close()


def test_e0_bvh_path_matches_numpy_oracle(hip_device):
    """Larger scene through the BVH traversal path (lexicographic ties):
    primary hits equal the oracle's brute-force scan on every pixel, and
    every pixel agrees within the tolerance (the BVH only orders the same
    candidate tests, so nothing may differ)."""
    scene, cam = workloads.terrain(4, nx=41, nz=21, chunk=20, emissive_quads=3)
    r = HipReferenceRenderer(samples_per_pixel=1, max_bounces=3, seed=2)
    img = r.render_frame(scene, cam, 80, 48, capture_primary_hits=True)
    trace = []
    ref = e0_numpy.render_scene(scene, cam, 80, 48, samples=1, max_bounces=3, seed=2, trace=trace)
    same = r.last_primary_hits == trace[0][3]
    ok = rel_close(img[..., :3], ref[..., :3])
    log(f"  e0 bvh within tol {ok.mean():.6f}, primary hits equal {same.mean():.6f}")
    assert same.all() and ok.all()
    r.close()


@pytest.mark.parametrize("case", ["room", "terrain"])
def test_named_products_match_oracle(hip_device, case):
    """depth / normal / primitive of sample 0's primary hit bit-exact with the
    oracle's traversal of the same camera rays; ids mapped like the reference;
    variance within 1 % on >= 99 % of pixels."""
    if case == "room":
        scene, cam = workloads.box_room(5)
        w, h, spp, bounces = 48, 36, 6, 4
    else:
        scene, cam = workloads.terrain(6, nx=141, nz=71, chunk=35, emissive_quads=6)
        w, h, spp, bounces = 64, 40, 5, 6
    outputs = ("color", "variance", "depth", "normal", "instance_id", "object_id", "material_id")
    r = HipPathTracer(max_bounces=bounces, samples_per_pixel=spp)
    got = r.render_products(scene, cam, w, h, outputs=outputs)
    color_only = r.render_frame(scene, cam, w, h)
    r.close()
    assert got["depth"].dtype == np.float32 and got["depth"].shape == (h, w)
    assert got["normal"].shape == (h, w, 3) and got["instance_id"].dtype == np.uint32
    assert np.array_equal(got["color"], color_only)
    orc = oracle.OracleScene(packing.pack_scene(scene))
    ref = orc.products(cameras.camera_basis(cam), cameras.camera_projection(cam), w, h, spp, 0, max_bounces=bounces)
    orc.close()
    tri_inst = scene.triangle_instance_ids()
    tri_mat = scene.triangle_material_ids()
    prim = ref["primitive"]
    hit = prim != 0xFFFFFFFF
    exp_inst = np.full((h, w), 0xFFFFFFFF, np.uint32)
    exp_inst[hit] = tri_inst[prim[hit]]
    exp_mat = np.full((h, w), 0xFFFFFFFF, np.uint32)
    exp_mat[hit] = tri_mat[prim[hit]]
    same = got["instance_id"] == exp_inst
    log(f"products {case}: hit {hit.mean():.3f} instance ids equal {same.mean():.5f}")
    # Both sides resolve the same camera rays by the same closest-hit rule
    # (lowest distance, then lowest triangle index), so every pixel agrees.
    assert same.all() and np.array_equal(got["instance_id"], got["object_id"])
    assert np.array_equal(got["material_id"], exp_mat)
    ok = same & (got["material_id"] == exp_mat)
    assert np.array_equal(got["depth"][ok].view(np.uint32), ref["depth"][ok].view(np.uint32))
    assert np.array_equal(got["normal"][ok].view(np.uint32), ref["normal"][ok].view(np.uint32))
    assert np.all(np.isinf(got["depth"][~hit & ok])) and np.all(got["normal"][~hit & ok] == 0)
    v, rv = got["variance"], ref["variance"]
    close = np.abs(v - rv) <= 1e-6 + 1e-2 * np.maximum(np.abs(v), np.abs(rv))
    log(f"  variance within 1 %: {close.mean():.5f}")
    assert close.mean() >= 0.99 and np.all(v >= 0)


@pytest.mark.parametrize("case", ["room", "lights"])
def test_gi_russian_roulette_matches_oracle(hip_device, case):
    """Russian roulette on (start 2, minimum survival 0.1 and 0.5): the
    survival draw, termination and 1/p_survive reweighting follow the same
    random stream on both sides (wavefront_shade_candidate.glsl:1335-1354)."""
    from ordinarylight_amd.scene import Material, Transform
    scene, cam = workloads.box_room(6)
    if case == "lights":
        glass = Material(base_color=(0.9, 0.95, 1.0), transmission=1.0, ior=1.5, roughness=0.0)
        v, i = workloads._box((0.4, 0.4, 0.4), None)
        scene.add_mesh(v, i, glass, transform=Transform.translation((0.1, 0.7, 0.3)))
        scene.add_point_light((0.3, 1.6, 0.4), (1.0, 0.9, 0.8), 2.0)
    for survival in (0.1, 0.5):
        log(f"gi roulette {case} min survival {survival}")
        img, ref, stats = gi_compare(scene, cam, 80, 60, 4, 8, wavefront_russian_roulette=True,
                                     wavefront_russian_roulette_start=2,
                                     wavefront_russian_roulette_min_survival=survival)
        ok = rel_close(img[..., :3], ref[..., :3])
        log(f"  within tol {ok.mean():.6f} extension rays {stats.extension_rays}")
        assert ok.mean() >= 0.999
    # Roulette ends paths early: fewer extension rays than without it.
    _, _, plain = gi_compare(scene, cam, 80, 60, 4, 8)
    assert stats.extension_rays < plain.extension_rays


def test_gi_sah_hierarchy_matches_oracle(hip_device):
    """The host binned-SAH hierarchy (the benchmark's builder) on a terrain
    with a material per triangle: same image and primary hits as the oracle."""
    scene, cam = workloads.terrain(7, nx=161, nz=81, chunk=40, emissive_quads=8)
    log("gi sah terrain")
    img, ref, stats = gi_compare(scene, cam, 128, 72, 2, 8, bvh_builder="sah")
    ok = rel_close(img[..., :3], ref[..., :3])
    log(f"  within tol {ok.mean():.6f} exact {np.all(img == ref, axis=-1).mean():.6f}")
    assert ok.mean() >= 0.999


def test_frame_index_selects_the_sample_stream(hip_device):
    """The image depends on frame_index only (the renderer keeps no frame
    counter of its own), and the index is stored as float32 like the
    reference's camera block (frames 2^24 and 2^24 + 1 share a stream)."""
    scene, cam = workloads.box_room(2)
    a = HipPathTracer(max_bounces=3, samples_per_pixel=1)
    frames = [a.render_frame(scene, cam, 32, 24, frame_index=k) for k in range(3)]
    again = a.render_frame(scene, cam, 32, 24, frame_index=1)
    big = [a.render_frame(scene, cam, 32, 24, frame_index=k) for k in (1 << 24, (1 << 24) + 1)]
    a.close()
    b = HipPathTracer(max_bounces=3, samples_per_pixel=1)
    fresh = b.render_frame(scene, cam, 32, 24, frame_index=2)
    b.close()
    assert np.array_equal(frames[1], again) and np.array_equal(frames[2], fresh)
    assert not np.array_equal(frames[0], frames[1]) and not np.array_equal(frames[1], frames[2])
    assert np.array_equal(big[0], big[1])


def test_e0_image_environment_matches_numpy_oracle(hip_device):
    """Nearest-texel image sky with rotation (path_tracer.py:184-202), seen
    directly by a panoramic camera and through the room's open side."""
    from ordinarylight_amd.lights import EnvironmentLight
    scene, cam = workloads.box_room(3)
    rng = np.random.default_rng(12)
    image = rng.uniform(0.0, 2.0, (16, 32, 3)).astype(np.float32)
    scene.set_environment(EnvironmentLight(image=image, color=(0.9, 1.0, 0.8), intensity=0.8, rotation=0.7))
    for kind, camera in (("pinhole", cam), ("pano", cameras.PanoramicCamera((0.0, 1.0, 0.5), (0.0, 1.0, 3.0)))):
        log(f"e0 image sky {kind}")
        r = HipReferenceRenderer(samples_per_pixel=2, max_bounces=3, seed=4)
        img = r.render_frame(scene, camera, 48, 40, frame_index=1)
        r.close()
        ref = e0_numpy.render_scene(scene, camera, 48, 40, samples=2, max_bounces=3, seed=4, frame_index=1)
        ok = rel_close(img[..., :3], ref[..., :3])
        log(f"  within tol {ok.mean():.6f} exact {np.all(img == ref, axis=-1).mean():.6f}")
        assert ok.mean() >= 0.999


def test_motion_product_matches_oracle_hits(hip_device):
    """The motion output over two frames of a static scene seen by a moving
    camera: zero on the first frame; on the second, the projection of each
    pixel's primary hit under the new camera minus its projection under the
    old one, with the hits taken from the oracle's traversal of the same
    camera rays (the device's hits equal them bit for bit)."""
    from ordinarylight_amd.cameras import PerspectiveCamera
    from ordinarylight_amd.renderers.hip import motion_vectors, capture_motion_state
    scene, cam0 = workloads.box_room(5)
    cam1 = PerspectiveCamera(position=tuple(np.asarray(cam0.position) + (0.05, 0.02, 0.0)),
                             target=tuple(np.asarray(cam0.target) + (0.02, 0.0, 0.0)), up=cam0.up,
                             vertical_fov_degrees=cam0.vertical_fov_degrees)
    w, h = 64, 48
    r = HipPathTracer(max_bounces=3, samples_per_pixel=2)
    first = r.render_products(scene, cam0, w, h, outputs=("motion", "color"), frame_index=0)["motion"]
    second = r.render_products(scene, cam1, w, h, outputs=("motion",), frame_index=1)["motion"]
    r.close()
    assert first.shape == (h, w, 2) and not first.any()
    orc = oracle.OracleScene(packing.pack_scene(scene))
    o, d = oracle.camera_rays(cameras.camera_basis(cam1), cameras.camera_projection(cam1), w, h, 1, 0)
    t, u, v, prim = orc.trace(o[:, :3], d[:, :3], float(o[0, 3]), float(d[0, 3]))
    orc.close()
    position = o[:, :3] + t[:, None] * d[:, :3]
    expect = motion_vectors(scene, cam1, w, h, prim, position, np.stack((u, v), axis=1),
                            capture_motion_state(scene, cam0, (w, h)))
    log(f"motion: max |m| {np.abs(second).max():.3f} px, max error {np.abs(second - expect).max():.2e}")
    assert np.abs(expect).max() > 0.5
    assert np.allclose(second, expect, atol=1e-4)
