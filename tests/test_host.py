"""Host-side checks: scene packing KATs, the C ABI's exports, the renderer contract.

* Packing known answers from the reference's own tests
  (tests/test_renderer.py:149-171 of the reference).
* Every symbol declared in include/olhip.h is exported by libolhip.so and
  bound by ordinarylight_amd._native (no GPU call is made).
* The implementation contract the reference's own ``Renderer`` relies on:
  call signatures, capabilities, validation and close semantics (the
  frame-index sequence itself is the caller's; its effect on the image is
  a GPU test).
"""

import ctypes
import inspect
import pathlib
import re

import numpy as np
import pytest

from ordinarylight_amd import cameras, packing, workloads
from ordinarylight_amd.scene import Material, Scene

ROOT = pathlib.Path(__file__).resolve().parent.parent


# -- packing -----------------------------------------------------------------------
def test_emissive_weights_by_area_and_power():
    scene = Scene()
    scene.add_mesh(((0, 0, 0), (1, 0, 0), (0, 2, 0)), ((0, 1, 2),), Material(emission=(1, 1, 1)))
    scene.add_mesh(((0, 0, 1), (2, 0, 1), (0, 2, 1)), ((0, 1, 2),), Material(emission=(2, 2, 2)))
    data = scene.emissive_triangle_data()
    np.testing.assert_allclose(data[:, 4, 1], (0.2, 0.8), rtol=1e-6)
    np.testing.assert_allclose(data[:, 4, 0], (0.2, 1.0), rtol=1e-6)
    assert abs(scene.emissive_light_weight - 5.0) < 1e-9
    assert scene.emissive_triangle_count == 2


def test_two_sided_emission_flag_is_packed():
    scene = Scene()
    scene.add_mesh(((0, 0, 0), (1, 0, 0), (0, 1, 0)), ((0, 1, 2),),
                   Material(emission=(1, 1, 1), emission_two_sided=True))
    assert scene.emissive_triangle_data()[0, 4, 2] == 1.0
    p = packing.pack_scene(scene)
    assert p.materials[0, 3, 3] == 1.0


def test_no_emitters_packs_placeholder_record():
    scene, _ = workloads.box_room(0)
    for mesh in scene.meshes:
        mesh.material = Material()
    assert scene.emissive_triangle_count == 0 and scene.emissive_light_weight == 0


def test_packing_order_ids_and_materials():
    scene, _ = workloads.box_room(0)
    p = packing.pack_scene(scene)
    tris, colors, emissions = scene.triangles()
    assert np.array_equal(p.positions[p.indices], tris)
    assert np.array_equal(p.materials[p.material_ids, 0, :3], colors)
    assert np.array_equal(p.materials[p.material_ids, 1, :3], emissions)
    # Occlusion factor resolved for untextured materials.
    assert np.all(p.materials[:, 5, 3] == 1.0)
    # IDs follow the packed triangle order.
    ids = scene.triangle_instance_ids()
    assert len(ids) == 36 and ids[0] == scene.meshes[0].id
    assert p.area_light_count == 2 and abs(p.area_light_weight - 0.25 * 15.0) < 1e-6


def test_analytic_light_records():
    scene = Scene()
    scene.add_point_light((1, 2, 3), (0.5, 0.5, 0.5), 4.0, range=10.0)
    scene.add_directional_light((0, -2, 0))
    scene.set_environment(color=(1, 1, 1), intensity=0.5)
    data = scene.analytic_light_data()
    assert data.shape == (3, 4, 4) and scene.analytic_light_count == 3
    assert tuple(data[0, 0]) == (1, 2, 3, 0) and data[0, 1, 3] == 10.0
    assert tuple(data[1, 1, :3]) == (0, -1, 0) and data[1, 0, 3] == 1
    assert data[2, 0, 3] == 3 and data[2, 2, 3] == 0.5 and data[2, 3, 0] == -1


def test_terrain_workload_shape():
    scene, cam = workloads.terrain(0, nx=101, nz=51, chunk=25, emissive_quads=32)
    assert scene.triangle_count == 100 * 50 * 2 + 64
    assert scene.emissive_triangle_count == 64
    assert scene.environment.intensity == 0.5
    faces = [m.face_materials for m in scene.meshes[:-1]]
    rough = np.concatenate([f.roughness for f in faces])
    metal = np.concatenate([f.metallic for f in faces])
    assert len(rough) == 100 * 50 * 2
    assert rough.min() >= 0.05 and rough.max() <= 1.0 and set(np.unique(metal)) == {0.0, 1.0}
    p = packing.pack_scene(scene)
    # Per-triangle parameters over the terrain chunks' shared records; the
    # expanded form has one full record per triangle.
    assert p.face_params.shape == (scene.triangle_count, 5)
    np.testing.assert_array_equal(p.face_params[:len(rough), 3], rough)
    np.testing.assert_array_equal(p.face_params[:len(rough), 4], metal)
    full, ids = packing.expanded_materials(p)
    assert len(full) == scene.triangle_count
    np.testing.assert_array_equal(full[ids[:len(rough)], 0, 3], rough)
    np.testing.assert_array_equal(full[ids[:len(rough)], 1, 3], metal)
    base = np.concatenate([f.base_color for f in faces])
    np.testing.assert_array_equal(full[:len(rough), 0, :3], base)
    # Triangles without face materials keep their record's values.
    emit = p.materials[p.material_ids[-1]]
    np.testing.assert_array_equal(full[-1], emit)
    chunked, _ = workloads.terrain(0, nx=101, nz=51, chunk=25, emissive_quads=32, materials="chunk")
    assert len(packing.pack_scene(chunked).materials) == 4 * 2 + 1


def test_full_terrain_has_one_million_triangles():
    scene, _ = workloads.terrain(0)
    assert scene.triangle_count - 64 == 1_000_000


def test_camera_basis_matches_reference_definition():
    cam = cameras.PerspectiveCamera((0, 1, 3), (0, 1, 0), vertical_fov_degrees=60)
    pos, fwd, right, up = cameras.camera_basis(cam)
    assert fwd.dtype == np.float32
    np.testing.assert_allclose(fwd, (0, 0, -1))
    np.testing.assert_allclose(right, (np.tan(np.radians(30)), 0, 0), rtol=1e-6)
    np.testing.assert_allclose(up, (0, np.tan(np.radians(30)), 0), rtol=1e-6)


# -- C ABI ---------------------------------------------------------------------------
def test_library_exports_every_declared_symbol():
    from ordinarylight_amd import _native
    header = (ROOT / "include" / "olhip.h").read_text()
    declared = set(re.findall(r"^\s*(?:int|const char\*)\s+(olhip_\w+)\s*\(", header, re.M))
    assert declared, "no declarations parsed"
    assert declared == set(_native.EXPORTED_SYMBOLS)
    lib = ctypes.CDLL(str(_native.LIBRARY_PATH))
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.olhip_version() == 100


def test_library_reports_errors_without_device():
    from ordinarylight_amd import _native
    lib = _native.library()
    count = ctypes.c_int(-1)
    assert lib.olhip_device_count(ctypes.byref(count)) == 0
    if count.value == 0:
        handle = ctypes.c_void_p()
        assert lib.olhip_create(0, ctypes.byref(handle)) == _native.OLHIP_ERR_RUNTIME
        assert b"no HIP device" in lib.olhip_last_error()
    assert lib.olhip_set_profiling(None, 1) == _native.OLHIP_ERR_VALUE


# -- renderer contract ------------------------------------------------------------------
def _keyword_only(fn):
    return {n: p.default for n, p in inspect.signature(fn).parameters.items() if p.kind is p.KEYWORD_ONLY}


def test_protocol_signatures_match_the_reference_contract():
    """The reference's Renderer calls implementations as
    render_frame(scene, camera, width, height, *, samples, frame_index) and
    render_products(..., *, outputs, samples, frame_index)
    (renderer/_core.py:757-766, renderers/base.py:41-71)."""
    from ordinarylight_amd.renderers.hip import HipPathTracer, HipReferenceRenderer
    for cls in (HipPathTracer, HipReferenceRenderer):
        params = list(inspect.signature(cls.render_frame).parameters)
        assert params[:5] == ["self", "scene", "camera", "width", "height"], cls
        kw = _keyword_only(cls.render_frame)
        assert kw["samples"] is None and kw["frame_index"] == 0, cls
    kw = _keyword_only(HipPathTracer.render_products)
    assert set(kw) >= {"outputs", "samples", "frame_index"}
    assert kw["samples"] is None and kw["frame_index"] == 0


def test_close_is_idempotent_and_final():
    from ordinarylight_amd.renderers.hip import HipPathTracer
    scene, cam = workloads.box_room(0)
    r = HipPathTracer()      # no device is touched before the first render
    r.close()
    r.close()
    for call in (lambda: r.render_frame(scene, cam, 4, 4), lambda: r.replace_scene(scene),
                 lambda: r.render_products(scene, cam, 4, 4, outputs=("color",))):
        with pytest.raises(RuntimeError):
            call()


def test_capabilities_are_a_plain_mapping():
    """The reference normalises a dict of capabilities itself
    (capabilities/_core.py:78-99): outputs and limits must be consistent."""
    from ordinarylight_amd.renderers.hip import HipPathTracer, HipReferenceRenderer
    for cls in (HipPathTracer, HipReferenceRenderer):
        r = cls()
        caps = r.capabilities
        assert isinstance(caps, dict)
        assert tuple(caps["outputs"]) == tuple(r.available_outputs)
        assert caps["renderer"] == r.implementation.name
        assert caps["limits"]["max_bounces"] == 64
        r.close()


def test_render_products_validates_outputs_before_rendering():
    from ordinarylight_amd.renderers.hip import HipPathTracer
    scene, cam = workloads.box_room(0)
    r = HipPathTracer()
    for bad in ((), ("color", "albedo"), ("albedo",), ("color", "color")):
        with pytest.raises(ValueError):
            r.render_products(scene, cam, 4, 4, outputs=bad)
    with pytest.raises(ValueError):
        r.render_products(scene, cam, 0, 4, outputs=("color",))
    r.close()


def test_hip_renderer_configuration_validation():
    from ordinarylight_amd.renderers.hip import HipPathTracer, HipReferenceConfig, HipRendererConfig
    with pytest.raises(ValueError):
        HipRendererConfig(max_bounces=0)
    with pytest.raises(ValueError):
        HipRendererConfig(samples_per_pixel=65)
    with pytest.raises(ValueError):
        HipRendererConfig(wavefront_environment_samples=5)
    with pytest.raises(ValueError):
        HipReferenceConfig(samples_per_pixel=0)
    with pytest.raises(TypeError):
        HipPathTracer(config=HipRendererConfig(), max_bounces=3)
    r = HipPathTracer(max_bounces=3)  # no device touched until a render
    assert r.capabilities["renderer"] == "hip-wavefront-gi"
    assert r.implementation.graphics_api == "hip"
    r.reconfigure(max_bounces=6)
    assert r.config.max_bounces == 6
    with pytest.raises(RuntimeError):
        r.reconfigure(device_index=1)
    r.close()
    with pytest.raises(RuntimeError):
        r.render_frame(None, None, 4, 4)
    from ordinarylight_amd.renderers import RendererProtocol
    assert isinstance(r, RendererProtocol)


def sah(boxes):
    from ordinarylight_amd import _native as N
    n = len(boxes)
    boxes = np.ascontiguousarray(boxes, np.float32)
    out = [np.zeros(n, np.uint32)] + [np.zeros(n - 1, np.uint32) for _ in range(4)]
    N.check(N.library().olhip_sah_topology(N.ptr(boxes), n, *[o.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
                                                               for o in out]))
    return out


def test_sah_top_level_hierarchy_is_a_full_binary_tree_over_contiguous_ranges():
    """The host SAH hierarchy of two-level scenes: every primitive sits in one
    leaf, every internal node covers the union of its children's contiguous
    ranges, and the tree is reachable from the root without cycles."""
    rng = np.random.default_rng(5)
    for n in (2, 3, 17, 1000, 300000):
        c = rng.uniform(-50, 50, (n, 3))
        h = rng.uniform(0.1, 3.0, (n, 3))
        boxes = np.concatenate([c - h, c + h], 1)
        if n == 17:
            boxes[:] = boxes[0]  # coincident centroids: median splits
        order, left, right, first, last = sah(boxes)
        if n == 300000:
            # Subtrees are built on several threads: the result must not depend on it.
            assert all(np.array_equal(a, b) for a, b in zip((order, left, right, first, last), sah(boxes)))
        assert sorted(order.tolist()) == list(range(n))
        LEAF = 0x80000000

        def span(ref):
            return (int(ref) & 0x7FFFFFFF, int(ref) & 0x7FFFFFFF) if ref & LEAF else (first[ref], last[ref])

        seen_leaves, seen_nodes, stack = [], set(), [0]
        while stack:
            i = stack.pop()
            assert i not in seen_nodes
            seen_nodes.add(i)
            (a0, a1), (b0, b1) = span(left[i]), span(right[i])
            assert a0 == first[i] and a1 + 1 == b0 and b1 == last[i]
            for ref in (left[i], right[i]):
                if ref & LEAF:
                    seen_leaves.append(int(ref) & 0x7FFFFFFF)
                else:
                    assert ref > i
                    stack.append(ref)
        assert seen_nodes == set(range(n - 1)) and sorted(seen_leaves) == list(range(n))


def test_sah_separates_two_clusters_at_the_root():
    """Two distant clusters: the root split is the gap between them."""
    rng = np.random.default_rng(6)
    a = rng.uniform(0, 1, (50, 3))
    b = rng.uniform(100, 101, (70, 3))
    c = np.concatenate([a, b])
    boxes = np.concatenate([c - 0.01, c + 0.01], 1)
    order, left, right, first, last = sah(boxes)
    lo, hi = (int(left[0]) & 0x7FFFFFFF, int(right[0]) & 0x7FFFFFFF)
    left_prims = order[first[lo]:last[lo] + 1] if not left[0] & 0x80000000 else order[lo:lo + 1]
    assert set(left_prims.tolist()) in (set(range(50)), set(range(50, 120)))


def test_sah_splits_off_a_box_spanning_the_field():
    """A box as large as the whole field (a mesh of scattered emitters, say)
    among 2,000 small ones: centroid bins alone would keep it with the small
    boxes round the centre on every level, each node on its path taking its
    size; the top-level build splits it off alone within two levels of the
    root."""
    rng = np.random.default_rng(8)
    c = rng.uniform(-50, 50, (2000, 3))
    boxes = np.concatenate([c - 0.5, c + 0.5], 1)
    boxes = np.concatenate([boxes, [[-50, -50, -50, 50, 50, 50]]])
    order, left, right, first, last = sah(boxes)
    LEAF = 0x80000000
    leaf_of = {int(order[k]): k for k in range(len(order))}
    depth, node, target = 0, 0, leaf_of[2000]
    while True:
        nxt = None
        for ref in (left[node], right[node]):
            if ref & LEAF:
                if int(ref) & 0x7FFFFFFF == target:
                    nxt = -1
            elif first[ref] <= target <= last[ref]:
                nxt = int(ref)
        assert nxt is not None
        depth += 1
        if nxt < 0:
            break
        node = nxt
    assert depth <= 2
    # Many nested boxes of slowly shrinking size, each spanning most of the
    # field: the binned splits price at least as low as splitting off the
    # largest, so no chain of single-box levels forms (one level per box would
    # be 512 deep).
    c = rng.uniform(-1, 1, (512, 3))
    h = 50.0 * 0.995 ** np.arange(512)[:, None]
    boxes = np.concatenate([c - h, c + h], 1)
    order, left, right, first, last = sah(boxes)

    def depth_of(ref):
        return 0 if ref & LEAF else 1 + max(depth_of(left[ref]), depth_of(right[ref]))
    assert depth_of(0) <= 40


def test_packed_material_arrays_hold_one_record_per_material():
    """Both packers hand the C ABI [M, 12, 4] records, M = the material count
    it reads (48 floats each), and every material index lies below M."""
    from ordinarylight_amd import packing
    inst, _ = workloads.instanced_spheres(3, meshes=2, subdivisions=1, instances=12, box=6.0, emitters=4,
                                          emitter_size=0.3)
    p = packing.pack_instanced_scene(inst)
    assert p.materials.ndim == 3 and p.materials.shape[1:] == (12, 4)
    assert int(p.instance_material.max()) < len(p.materials)
    flat = packing.pack_scene(workloads.box_room(0)[0])
    assert flat.materials.ndim == 3 and flat.materials.shape[1:] == (12, 4)
    assert int(flat.material_ids.max()) < len(flat.materials)


def test_textured_scene_packing():
    """Textured materials: binding indices in record rows 4 / 5, the textured
    flag word, per-vertex uvs and world tangents (unit, handedness +-1), and
    an image environment's texture index in its analytic record."""
    from ordinarylight_amd import packing
    scene, _ = workloads.textured_room(0, open_sky=True)
    p = packing.pack_scene(scene)
    m = p.materials
    assert p.texture_words[0] == len(p.texture_words[1::8][:p.texture_words[0]])
    flags = m[:, 8, 3]
    assert set(flags.tolist()) <= {0.0, 1.0} and flags.sum() >= 2
    textured = m[flags == 1.0]
    for r in textured:
        for b in (*r[4], r[5, 1], r[5, 3]):
            assert b == -1.0 or 0 <= b < len(p.bindings)
    assert p.uvs.shape == (len(p.positions), 4) and p.tangents.shape == (len(p.positions), 4)
    assert np.allclose(np.linalg.norm(p.tangents[:, :3], axis=1), 1.0, atol=1e-5)
    assert set(np.unique(p.tangents[:, 3]).tolist()) <= {-1.0, 1.0}
    env = [r for r in p.analytic_lights[:p.analytic_light_count] if int(r[0, 3] + 0.5) == 3]
    assert len(env) == 1 and env[0][3, 0] == p.texture_words[0] - 1 and env[0][3, 2] > 0.0
    with pytest.raises(ValueError):
        packing.pack_instanced_scene(scene)


def test_motion_vectors_closed_forms():
    """The motion output's host math (the reference's opt-in product):
    zero without history, zero for a static scene and camera, the
    projection difference for a moved camera, the mesh's own displacement
    for a moved mesh, and zero where a mesh's triangles changed."""
    from ordinarylight_amd import workloads
    from ordinarylight_amd.cameras import PerspectiveCamera
    from ordinarylight_amd.renderers.hip import capture_motion_state, motion_vectors, project_positions
    from ordinarylight_amd.scene import Transform
    scene, cam = workloads.box_room(0)
    w, h = 16, 12
    packed = packing.pack_scene(scene)
    rng = np.random.default_rng(3)
    tri = rng.integers(0, len(packed.indices), (h, w)).astype(np.uint32)
    bary = rng.uniform(0, 0.5, (h, w, 2)).astype(np.float32)
    corners = packed.positions[packed.indices[tri.reshape(-1)]].astype(np.float64)
    weights = np.stack((1 - bary[..., 0].reshape(-1) - bary[..., 1].reshape(-1), bary[..., 0].reshape(-1),
                        bary[..., 1].reshape(-1)), 1)
    position = np.einsum("nk,nkc->nc", weights, corners).astype(np.float32).reshape(h, w, 3)
    assert not motion_vectors(scene, cam, w, h, tri, position, bary, None).any()
    state = capture_motion_state(scene, cam, (w, h))
    still = motion_vectors(scene, cam, w, h, tri, position, bary, state)
    assert np.abs(still).max() < 1e-3
    moved = PerspectiveCamera(position=tuple(np.asarray(cam.position) + (0.1, 0.05, 0.0)),
                              target=tuple(np.asarray(cam.target) + (0.1, 0.05, 0.0)), up=cam.up,
                              vertical_fov_degrees=cam.vertical_fov_degrees)
    m = motion_vectors(scene, moved, w, h, tri, position, bary, state)
    now, ok1 = project_positions(position.reshape(-1, 3), moved, w, h)
    then, ok2 = project_positions(position.reshape(-1, 3), cam, w, h)
    expect = np.where((ok1 & ok2)[:, None], now - then, 0.0).reshape(h, w, 2)
    assert np.allclose(m, expect, atol=1e-3) and np.abs(expect).max() > 0.1
    # A mesh moved by a translation: its pixels move by the projection of
    # the displacement, other meshes' pixels do not move.
    mesh = scene.visible_meshes[0]
    first = 0
    count = len(mesh.indices)
    old = mesh.transform.matrix.copy()
    mesh.transform = Transform(np.asarray(old) @ np.array(
        [[1, 0, 0, 0.2], [0, 1, 0, 0.0], [0, 0, 1, 0.0], [0, 0, 0, 1.0]]))
    shifted = position.copy()
    on_mesh = (tri >= first) & (tri < first + count)
    shifted[on_mesh] += np.float32(0.2) * np.asarray(old)[:3, 0].astype(np.float32)
    m2 = motion_vectors(scene, cam, w, h, tri, shifted, bary, state)
    now, a = project_positions(shifted.reshape(-1, 3), cam, w, h)
    then, b = project_positions(position.reshape(-1, 3), cam, w, h)
    exp2 = np.where((a & b)[:, None], now - then, 0.0).reshape(h, w, 2)
    assert on_mesh.any() and np.allclose(m2[on_mesh], exp2[on_mesh], atol=2e-3)
    assert np.abs(m2[~on_mesh]).max() < 1e-3
    # Changed triangles: no motion for that mesh.
    state2 = capture_motion_state(scene, cam, (w, h))
    key = mesh.id
    v, i, t = state2["meshes"][key]
    state2["meshes"][key] = (v, i[::-1].copy(), t)
    m3 = motion_vectors(scene, moved, w, h, tri, shifted, bary, state2)
    assert not m3[on_mesh].any()
