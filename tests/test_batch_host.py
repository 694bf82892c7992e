"""CPU: host logic of the batched/served workloads (no device calls)."""

import numpy as np
import pytest

from ordinarylight_amd import workloads
from ordinarylight_amd.batch import EngineSettings, autotune, latency_percentiles, poisson_arrivals


def test_poisson_arrivals_rate_and_determinism():
    a = poisson_arrivals(50.0, 20000, seed=3)
    assert np.all(np.diff(a) > 0)
    assert abs(len(a) / a[-1] - 50.0) < 1.5  # mean rate within ~3 %
    assert np.array_equal(a, poisson_arrivals(50.0, 20000, seed=3))
    # Exponential gaps: the coefficient of variation of the gaps is 1.
    gaps = np.diff(a)
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.05
    with pytest.raises(ValueError):
        poisson_arrivals(0.0, 3)


def test_latency_percentiles_nearest_rank():
    lat = list(range(1, 101))  # 1..100 ms
    p = latency_percentiles(lat)
    assert p["p50_ms"] == 50.0 and p["p99_ms"] == 99.0 and p["count"] == 100
    assert latency_percentiles([7.0])["p99_ms"] == 7.0
    assert latency_percentiles([3.0, 1.0, 2.0])["p50_ms"] == 2.0
    with pytest.raises(ValueError):
        latency_percentiles([])


def test_autotune_coordinate_search_picks_minimum():
    # Synthetic cost: best shape (8, 128), best capacity 2^23.
    def cost(s):
        shape = {(16, 128): 1.0, (8, 128): 0.8, (16, 256): 1.2, (32, 128): 1.1}[(s.lds_stack, s.trace_block)]
        size = {1 << 22: 1.3, 1 << 23: 1.0, 1 << 24: 1.05, 1 << 25: 1.1, 1 << 26: 1.15, 1 << 27: 1.2}[s.path_slots]
        return shape * size

    best, timings = autotune(None, [object()], measure=cost)
    assert (best.lds_stack, best.trace_block, best.path_slots) == (8, 128, 1 << 23)
    # Shape search at the largest capacity (4) + capacity search at the best shape (5 new).
    assert len(timings) == 9
    assert all(isinstance(k, EngineSettings) for k in timings)


def test_mixed_jobs_distribution():
    jobs = workloads.mixed_jobs(7, 400)
    assert jobs == workloads.mixed_jobs(7, 400)
    edges = np.array([j["width"] for j in jobs])
    assert all(j["width"] == j["height"] for j in jobs)
    assert edges.min() >= 128 and edges.max() <= 1024
    # Log-uniform edge: log(edge) roughly uniform on [log 128, log 1024].
    u = (np.log(edges) - np.log(128)) / (np.log(1024) - np.log(128))
    assert abs(u.mean() - 0.5) < 0.05
    assert {j["samples"] for j in jobs} == {1, 2, 4, 8, 16}
    assert {j["max_bounces"] for j in jobs} == {4, 8}
    assert {j["scene"] for j in jobs} == {k for k, _, _ in workloads.MIXED_SCENES}


def test_tessellated_room_size_and_shared_layout():
    scene, cam = workloads.tessellated_room(0, 10_000)
    n = scene.triangle_count
    assert 9_000 <= n <= 11_000
    base, _ = workloads.box_room(0)
    # Same boxes and light as the 36-triangle room: its last 26 triangles.
    assert np.allclose(scene.triangles()[0][-26:], base.triangles()[0][-26:], atol=1e-6)
    # Walls cover the same planes: every wall vertex lies on the room's box.
    v = scene.triangles()[0][:-26].reshape(-1, 3)
    on_box = (np.isclose(np.abs(v[:, 0]), 1) | np.isclose(v[:, 1], 0) | np.isclose(v[:, 1], 2)
              | np.isclose(np.abs(v[:, 2]), 1))
    assert on_box.all()


def test_instanced_packing_bounds_contain_instances_and_flatten_in_order():
    """World bounds of the two-level packing contain every transformed vertex
    and are no larger than the transformed object-box corners; triangle
    offsets follow the flattened packed order."""
    from ordinarylight_amd import packing
    scene, _ = workloads.instanced_spheres(3, meshes=3, subdivisions=2, instances=120, box=12.0, emitters=10)
    p = packing.pack_instanced_scene(scene)
    meshes = [m for m in scene.meshes if len(m.indices)]
    assert len(p.instance_resource) == len(meshes)
    offset = 0
    for k, mesh in enumerate(meshes):
        w = mesh.world_vertices
        lo, hi = p.world_bounds[k, :3], p.world_bounds[k, 3:]
        assert np.all(w >= lo) and np.all(w <= hi)
        corners = np.array([[(mesh.vertices.max(0) if b else mesh.vertices.min(0))[a] for a, b in enumerate(bits)]
                            for bits in np.ndindex(2, 2, 2)], np.float64)
        world = corners @ mesh.transform.matrix[:3, :3].T + mesh.transform.matrix[:3, 3]
        assert np.all(lo >= np.nextafter(world.min(0).astype(np.float32), -np.inf))
        assert np.all(hi <= np.nextafter(world.max(0).astype(np.float32), np.inf))
        assert p.instance_triangle_offset[k] == offset
        offset += len(mesh.indices)


def test_instanced_packing_spheres_contain_instances():
    """Each instance's world bounding sphere (float32 centre and radius)
    contains every transformed vertex and is tight: its radius exceeds the
    largest vertex distance from the centre by a relative 1e-4 at most."""
    from ordinarylight_amd import packing
    scene, _ = workloads.instanced_spheres(5, meshes=3, subdivisions=2, instances=200, box=12.0, emitters=10)
    p = packing.pack_instanced_scene(scene)
    meshes = [m for m in scene.meshes if len(m.indices)]
    assert p.world_spheres.shape == (len(meshes), 4) and p.world_spheres.dtype == np.float32
    for k, mesh in enumerate(meshes):
        m = np.asarray(mesh.transform.matrix, np.float64)
        w = mesh.vertices.astype(np.float64) @ m[:3, :3].T + m[:3, 3]
        c = p.world_spheres[k, :3].astype(np.float64)
        r = float(p.world_spheres[k, 3])
        far = np.sqrt(((w - c) ** 2).sum(1)).max()
        assert far <= r
        assert r <= far * (1.0 + 1e-4) + 1e-6


def test_execution_strategy_settings():
    """Strategies: wavefront (default) and hybrid reach the C ABI as the
    fused-primary flag; the reference's other strategies are refused."""
    from ordinarylight_amd.renderers.hip import HipRendererConfig
    assert not HipRendererConfig().fused_primary
    assert HipRendererConfig(wavefront_execution_strategy="hybrid").fused_primary
    assert not HipRendererConfig(wavefront_execution_strategy="auto").fused_primary
    for name in ("megakernel", "persistent", "ser", "fused"):
        with pytest.raises(ValueError):
            HipRendererConfig(wavefront_execution_strategy=name)
    assert EngineSettings().native().fused_primary == 0
    assert EngineSettings(fused_primary=True).native().fused_primary == 1
