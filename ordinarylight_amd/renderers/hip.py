"""``renderers.hip``: MI355X implementations of ``RendererProtocol``.

* :class:`HipPathTracer` -- the GI estimator of the reference's Vulkan
  wavefront (``VulkanGlobalIlluminationRenderer``,
  ``ordinarylight/renderers/gi/vulkan.py:11-20`` ->
  ``targets/vulkan/api.py:1233-1451``): same ``(H, W, 4)`` float32 linear
  HDR (alpha 1), same per-pixel hash sampling keyed on
  ``(pixel, frame_index, sample)``, same settings and limits.
* :class:`HipReferenceRenderer` -- the diffuse estimator of
  ``CpuReferenceRenderer`` (``renderers/reference/cpu.py:24-74``) with its
  sequential PCG64 stream seeded ``seed + frame_index`` (alpha 0).

Both drive ``libolhip.so`` through :mod:`ordinarylight_amd._native`; there
is no other compute path.
"""

from __future__ import annotations

from dataclasses import dataclass, fields, replace
import math
import threading
import time

import numpy as np

from .. import _native as N
from ..cameras import camera_basis, camera_projection
from ..packing import has_textured_materials, pack_instanced_scene, pack_scene, shares_resources
from ..selection import PickOptions, PickResult, ViewportMapping
from .base import RendererImplementation, RendererImplementationInfo


# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class HipRendererConfig:
    """Settings of the GI estimator (subset of ``RendererConfig``,
    ``targets/vulkan/api.py:26-184``, same names and defaults)."""

    max_bounces: int = 5
    samples_per_pixel: int = 1
    area_light_samples: int = 1
    wavefront_secondary_area_light_samples: int = 0
    wavefront_environment_samples: int = 1
    wavefront_secondary_nee_probability: float = 1.0
    wavefront_unified_secondary_nee: bool = True
    wavefront_russian_roulette: bool = False
    wavefront_russian_roulette_start: int = 3
    wavefront_russian_roulette_min_survival: float = 0.1
    # ReSTIR DI for area lights at primary hits, with the reference's
    # settings and defaults (targets/vulkan/api.py:60-75, limits :299-323);
    # candidates here go up to 64 (the reference stops at 4).  A job's
    # samples run as successive one-sample passes, each reusing the
    # previous pass's reservoirs.
    wavefront_restir_di: bool = False
    wavefront_restir_candidates: int = 1
    wavefront_restir_history_limit: int = 20
    wavefront_restir_spatial_reuse: bool = False
    wavefront_restir_spatial_neighbors: int = 4
    wavefront_restir_spatial_radius: int = 4
    wavefront_restir_pairwise_mis: bool = True
    # Exposure of the display outputs (api.py:107).
    wavefront_exposure: float = 1.0
    # Resident paths per device chunk (the reference's per-tile capacity is
    # 131,072; one MI355X holds whole frames, so the default is 2^25).
    wavefront_tile_capacity: int = 1 << 25
    # The reference accepts 1..64 samples per call (api.py:1270-1272).
    max_samples_per_call: int = 64
    device_index: int = 0
    bvh_leaf_size: int = 2
    # Bottom-level hierarchy: "lbvh" (GPU Morton LBVH) or "sah" (host binned
    # SAH, slower to build, fewer nodes visited per ray).
    bvh_builder: str = "lbvh"
    # Traversal stack entries kept in LDS per lane, and the traversal block size.
    traversal_lds_stack: int = 16
    traversal_block: int = 128
    profiling: bool = False
    work_counters: bool = False
    # Execution strategy (targets/vulkan/api.py:158, :611-616): "wavefront"
    # traces and shades every bounce in separate kernels; "hybrid" runs the
    # coherent primary bounce (fresh camera rays of flat, untextured scenes
    # without ReSTIR) as one traversal-and-shading kernel and later bounces
    # through the queues.  Both give the same image bit for bit; "auto" is
    # "wavefront", the faster of the two on the measured workloads.  The
    # reference's "megakernel", "persistent" and "ser" strategies are not
    # built.
    wavefront_execution_strategy: str = "wavefront"

    def __post_init__(self):
        if not 1 <= int(self.max_bounces) <= 64:
            raise ValueError("max_bounces must be between 1 and 64")
        if not 1 <= int(self.samples_per_pixel) <= int(self.max_samples_per_call):
            raise ValueError("samples_per_pixel must be between 1 and max_samples_per_call")
        if int(self.area_light_samples) < 1 or int(self.wavefront_secondary_area_light_samples) < 0:
            raise ValueError("area light sample counts must be positive")
        if not 0 <= int(self.wavefront_environment_samples) <= 4:
            raise ValueError("wavefront_environment_samples must be between 0 and 4")
        if not 0.0 < float(self.wavefront_secondary_nee_probability) <= 1.0:
            raise ValueError("wavefront_secondary_nee_probability must be in (0, 1]")
        if not 0.0 <= float(self.wavefront_russian_roulette_min_survival) <= 1.0:
            raise ValueError("wavefront_russian_roulette_min_survival must be in [0, 1]")
        if int(self.wavefront_tile_capacity) < 1:
            raise ValueError("wavefront_tile_capacity must be positive")
        if not 0 <= int(self.bvh_leaf_size) <= 16:
            raise ValueError("bvh_leaf_size must be between 0 and 16")
        if self.bvh_builder not in ("lbvh", "sah"):
            raise ValueError("bvh_builder must be 'lbvh' or 'sah'")
        if not (math.isfinite(float(self.wavefront_exposure)) and float(self.wavefront_exposure) > 0.0):
            raise ValueError("wavefront_exposure must be positive")
        if not 1 <= int(self.wavefront_restir_history_limit) <= 64:
            raise ValueError("wavefront_restir_history_limit must be between 1 and 64")
        if not 1 <= int(self.wavefront_restir_candidates) <= 64:
            raise ValueError("wavefront_restir_candidates must be between 1 and 64")
        for name in ("wavefront_restir_di", "wavefront_restir_spatial_reuse", "wavefront_restir_pairwise_mis"):
            if not isinstance(getattr(self, name), bool):
                raise TypeError(f"{name} must be a bool")
        if not 1 <= int(self.wavefront_restir_spatial_neighbors) <= 8:
            raise ValueError("wavefront_restir_spatial_neighbors must be between 1 and 8")
        if not 1 <= int(self.wavefront_restir_spatial_radius) <= 32:
            raise ValueError("wavefront_restir_spatial_radius must be between 1 and 32")
        if self.wavefront_execution_strategy not in ("wavefront", "hybrid", "auto"):
            raise ValueError("wavefront_execution_strategy must be 'wavefront', 'hybrid' or 'auto' "
                             "(the megakernel, persistent and ser strategies are not built)")

    @property
    def fused_primary(self):
        return self.wavefront_execution_strategy == "hybrid"

    def gi_settings(self):
        s = N.GiSettings()
        s.max_bounces = int(self.max_bounces)
        s.secondary_area_samples = int(self.wavefront_secondary_area_light_samples or self.area_light_samples)
        s.environment_samples = int(self.wavefront_environment_samples)
        s.russian_roulette_start = int(self.wavefront_russian_roulette_start) if self.wavefront_russian_roulette else 0
        s.russian_roulette_min_survival = float(self.wavefront_russian_roulette_min_survival)
        s.secondary_nee_probability = float(self.wavefront_secondary_nee_probability)
        s.unified_secondary_nee = int(bool(self.wavefront_unified_secondary_nee))
        s.path_capacity = int(min(self.wavefront_tile_capacity, 0xFFFFFFFF))
        if self.wavefront_restir_di:
            s.restir_candidates = int(self.wavefront_restir_candidates)
            s.restir_spatial_neighbors = int(self.wavefront_restir_spatial_neighbors) if self.wavefront_restir_spatial_reuse else 0
            s.restir_spatial_radius = int(self.wavefront_restir_spatial_radius)
            s.restir_history_limit = int(self.wavefront_restir_history_limit)
            s.restir_pairwise_mis = int(self.wavefront_restir_pairwise_mis)
        return s


# Settings that can change between frames without rebuilding device state
# (the reference's _HOT_SETTINGS, targets/vulkan/api.py:989-1043).
HOT_SETTINGS = frozenset({
    "max_bounces", "samples_per_pixel", "area_light_samples",
    "wavefront_secondary_area_light_samples", "wavefront_environment_samples",
    "wavefront_secondary_nee_probability", "wavefront_unified_secondary_nee",
    "wavefront_russian_roulette", "wavefront_russian_roulette_start",
    "wavefront_russian_roulette_min_survival", "wavefront_tile_capacity",
    "max_samples_per_call", "traversal_lds_stack", "traversal_block", "profiling", "work_counters",
    "wavefront_execution_strategy",
    "wavefront_exposure", "wavefront_restir_di", "wavefront_restir_candidates", "wavefront_restir_history_limit",
    "wavefront_restir_spatial_reuse", "wavefront_restir_spatial_neighbors", "wavefront_restir_spatial_radius",
    "wavefront_restir_pairwise_mis",
})

OUTPUT_FORMATS = {"rgba8": 0, "nv12": 1, "p010": 2}


class DeviceContext:
    """One ``olhip_ctx``: HIP stream and frame buffers on one device."""

    def __init__(self, device_index=0):
        self._lib = N.library()
        handle = N.ctypes.c_void_p()
        N.check(self._lib.olhip_create(int(device_index), N.ctypes.byref(handle)))
        self.handle = handle
        self.device_index = int(device_index)

    def camera_rays(self, camera, width, height, frame_index=0, sample=0, tile=None):
        """Primary rays of a pixel tile ``(x0, y0, w, h)`` (default: the whole
        image) exactly as a render generates them: float32 origins
        ``[n, 4]`` (xyz, tmin) and directions ``[n, 4]`` (xyz, tmax), row-major."""
        x0, y0, tw, th = tile if tile is not None else (0, 0, width, height)
        position, forward, right, up = camera_basis(camera)
        cam = N.Camera()
        cam.position[:] = position.tolist()
        cam.forward[:] = forward.tolist()
        cam.right[:] = right.tolist()
        cam.up[:] = up.tolist()
        cam.projection = camera_projection(camera)
        o = np.empty((tw * th, 4), np.float32)
        d = np.empty((tw * th, 4), np.float32)
        N.check(N.library().olhip_camera_rays(self.handle, N.ctypes.byref(cam), int(width), int(height),
                                              int(frame_index), int(sample), int(x0), int(y0), int(tw), int(th),
                                              N.ptr(o), N.ptr(d)))
        return o, d

    def close(self):
        if self.handle:
            N.check(self._lib.olhip_destroy(self.handle))
            self.handle = None


BUILDERS = {"lbvh": N.BUILDER_LBVH, "sah": N.BUILDER_SAH}


def _scene_desc(packed, leaf_size, builder="lbvh"):
    desc = N.SceneDesc()
    desc.vertex_count = len(packed.positions)
    desc.positions = N.ptr(packed.positions)
    desc.normals = N.ptr(packed.normals)
    desc.triangle_count = len(packed.indices)
    desc.indices = N.ptr(packed.indices, N.ctypes.c_uint32)
    desc.material_ids = N.ptr(packed.material_ids, N.ctypes.c_uint32)
    desc.material_count = len(packed.materials)
    desc.materials = N.ptr(packed.materials)
    desc.analytic_light_count = packed.analytic_light_count
    desc.analytic_lights = N.ptr(packed.analytic_lights)
    desc.area_light_count = packed.area_light_count
    desc.area_lights = N.ptr(packed.area_lights)
    desc.area_light_weight = packed.area_light_weight
    desc.leaf_size = int(leaf_size)
    desc.builder = BUILDERS[builder]
    if packed.texture_words is not None:
        desc.texture_word_count = len(packed.texture_words)
        desc.texture_words = N.ptr(packed.texture_words, N.ctypes.c_uint32)
    if packed.bindings is not None:
        desc.binding_count = len(packed.bindings)
        desc.bindings = N.ptr(packed.bindings)
    if packed.uvs is not None:
        desc.uvs = N.ptr(packed.uvs)
        desc.tangents = N.ptr(packed.tangents)
        desc.uv_density = N.ptr(packed.uv_density)
    if packed.face_params is not None:
        desc.face_params = N.ptr(packed.face_params)
    return desc


def _instanced_desc(p, leaf_size, builder="lbvh"):
    d = N.InstancedDesc()
    d.resource_count = len(p.vertex_offsets) - 1
    d.vertex_offsets = N.ptr(p.vertex_offsets, N.ctypes.c_uint32)
    d.triangle_offsets = N.ptr(p.triangle_offsets, N.ctypes.c_uint32)
    d.positions = N.ptr(p.positions)
    d.normals = N.ptr(p.normals)
    d.indices = N.ptr(p.indices, N.ctypes.c_uint32)
    d.instance_count = len(p.instance_resource)
    d.instance_resource = N.ptr(p.instance_resource, N.ctypes.c_uint32)
    d.instance_material = N.ptr(p.instance_material, N.ctypes.c_uint32)
    d.instance_triangle_offset = N.ptr(p.instance_triangle_offset, N.ctypes.c_uint32)
    d.instance_id = N.ptr(p.instance_id, N.ctypes.c_uint32)
    d.world_to_object = N.ptr(p.world_to_object)
    d.normal_matrix = N.ptr(p.normal_matrix)
    d.determinant = N.ptr(p.determinant)
    d.world_bounds = N.ptr(p.world_bounds)
    d.world_spheres = N.ptr(p.world_spheres)
    d.material_count = len(p.materials)
    d.materials = N.ptr(p.materials)
    d.analytic_light_count = p.analytic_light_count
    d.analytic_lights = N.ptr(p.analytic_lights)
    d.area_light_count = p.area_light_count
    d.area_lights = N.ptr(p.area_lights)
    d.area_light_weight = p.area_light_weight
    d.leaf_size = int(leaf_size)
    d.builder = BUILDERS[builder]
    if p.texture_words is not None:
        d.texture_word_count = len(p.texture_words)
        d.texture_words = N.ptr(p.texture_words, N.ctypes.c_uint32)
    return d


def _same(a, b):
    return a.shape == b.shape and np.array_equal(a, b)


class DeviceScene:
    """A scene resident in HBM: packed arrays + the GPU-built BVH.

    ``instanced``: True builds a two-level scene (a bottom-level BVH per mesh
    resource, a top-level BVH over instances), False flattens every instance
    into world space; None chooses two levels when a resource is placed more
    than once.
    """

    def __init__(self, context, scene, leaf_size=4, instanced=None, builder="lbvh"):
        self._lib = N.library()
        self.context = context
        self.leaf_size = int(leaf_size)
        if builder not in BUILDERS:
            raise ValueError("builder must be 'lbvh' or 'sah'")
        self.builder = builder
        if instanced is None:
            # Textured materials are shaded on flat scenes only.
            instanced = shares_resources(scene) and not has_textured_materials(scene)
        self.instanced = bool(instanced)
        if self.instanced:
            self._init_instanced(scene, leaf_size)
            return
        started = time.perf_counter()
        packed = pack_scene(scene)
        self.packed = packed
        pack_ms = (time.perf_counter() - started) * 1000.0
        desc = _scene_desc(packed, leaf_size, builder)
        handle = N.ctypes.c_void_p()
        started = time.perf_counter()
        N.check(self._lib.olhip_scene_create(context.handle, N.ctypes.byref(desc), N.ctypes.byref(handle)))
        self.handle = handle
        upload_ms = (time.perf_counter() - started) * 1000.0
        info = N.SceneInfo()
        N.check(self._lib.olhip_scene_info_get(handle, N.ctypes.byref(info)))
        self.info = info
        self.timings = {"scene_pack_ms": pack_ms, "scene_upload_ms": upload_ms, "bvh_ms": info.build_ms}

    def _init_instanced(self, scene, leaf_size):
        started = time.perf_counter()
        p = pack_instanced_scene(scene)
        self.packed = p
        pack_ms = (time.perf_counter() - started) * 1000.0
        d = _instanced_desc(p, leaf_size, self.builder)
        handle = N.ctypes.c_void_p()
        started = time.perf_counter()
        N.check(self._lib.olhip_scene_create_instanced(self.context.handle, N.ctypes.byref(d), N.ctypes.byref(handle)))
        self.handle = handle
        upload_ms = (time.perf_counter() - started) * 1000.0
        info = N.SceneInfo()
        N.check(self._lib.olhip_scene_info_get(handle, N.ctypes.byref(info)))
        self.info = info
        self.timings = {"scene_pack_ms": pack_ms, "scene_upload_ms": upload_ms, "bvh_ms": info.build_ms}

    def update(self, scene):
        """Bring the resident copy up to date with ``scene`` in place when its
        triangles are unchanged (the reference's partial update,
        ``targets/vulkan/core.py:4696-4870``): moved vertices refit the BVH,
        moved instances rebuild the top level, materials and lights are
        replaced.  Returns False, changing nothing, when the structure
        differs and the scene must be rebuilt."""
        started = time.perf_counter()
        old = self.packed
        if self.instanced:
            if not shares_resources(scene):
                return False
            p = pack_instanced_scene(scene)
            keys = ("vertex_offsets", "triangle_offsets", "indices", "instance_resource", "instance_triangle_offset")
            if not all(_same(getattr(p, k), getattr(old, k)) for k in keys):
                return False
            d = _instanced_desc(p, self.leaf_size, self.builder)
            update = self._lib.olhip_scene_update_instanced
        else:
            p = pack_scene(scene)
            if not (_same(p.indices, old.indices) and len(p.positions) == len(old.positions)):
                return False
            d = _scene_desc(p, self.leaf_size, self.builder)
            update = self._lib.olhip_scene_update
        moved = not (_same(p.positions, old.positions) and _same(p.normals, old.normals))
        if not moved:
            d.positions = None
            d.normals = None
        pack_ms = (time.perf_counter() - started) * 1000.0
        started = time.perf_counter()
        N.check(update(self.context.handle, self.handle, N.ctypes.byref(d)))
        self.packed = p
        info = N.SceneInfo()
        N.check(self._lib.olhip_scene_info_get(self.handle, N.ctypes.byref(info)))
        self.info = info
        self.timings = {"scene_pack_ms": pack_ms, "scene_update_ms": (time.perf_counter() - started) * 1000.0,
                        "bvh_ms": info.build_ms if moved else 0.0, "bvh_refit": bool(moved)}
        return True

    @property
    def triangle_count(self):
        return int(self.info.triangle_count)

    def bvh_arrays(self):
        """(nodes uint8 [N, 64], triangles float32 [T, 3, 4]) copied from the device."""
        nodes = np.zeros((self.info.node_count, 64), np.uint8)
        tri = np.zeros((self.info.triangle_count, 3, 4), np.float32)
        N.check(self._lib.olhip_scene_bvh(self.handle, nodes.ctypes.data_as(N.ctypes.c_void_p),
                                          tri.ctypes.data_as(N.ctypes.c_void_p)))
        return nodes, tri

    def trace(self, origins, directions, tmin=0.0, tmax=1e30):
        """Closest hits: returns (t float32, u, v, triangle uint32; 0xFFFFFFFF = miss)."""
        n = len(origins)
        o = np.empty((n, 4), np.float32)
        d = np.empty((n, 4), np.float32)
        o[:, :3] = origins
        o[:, 3] = tmin
        d[:, :3] = directions
        d[:, 3] = tmax
        hits = np.empty((n, 4), np.float32)
        N.check(self._lib.olhip_trace(self.context.handle, self.handle, N.ptr(o), N.ptr(d), n, N.ptr(hits)))
        return hits[:, 0], hits[:, 1], hits[:, 2], hits[:, 3].view(np.uint32)

    def close(self):
        if self.handle:
            N.check(self._lib.olhip_scene_destroy(self.handle))
            self.handle = None


class _ResidentMixin:
    """Scene residency keyed by identity and revision (``core.py:4574-4634``)."""

    def _init_resident(self, device_index, leaf_size, builder="lbvh"):
        self._context = None
        self._device_index = int(device_index)
        self._leaf_size = int(leaf_size)
        self._builder = builder
        self._scene_key = None
        self._scene_ref = None
        self._device_scene = None
        self._lock = threading.RLock()
        self._closed = False

    def _ensure_context(self):
        if self._context is None:
            self._context = DeviceContext(self._device_index)
        return self._context

    def _resident(self, scene):
        key = (id(scene), getattr(scene, "revision", None))
        if (self._device_scene is not None and self._scene_key != key and self._scene_key[0] == key[0]
                and self._scene_ref is scene and hasattr(self._device_scene, "update")
                and self._device_scene.update(scene)):
            # Same scene object, same triangles: updated in place.
            self._scene_key = key
            return self._device_scene
        if self._device_scene is None or self._scene_key != key:
            previous = self._device_scene
            self._device_scene = DeviceScene(self._ensure_context(), scene, self._leaf_size,
                                             builder=self._builder)
            self._scene_key = key
            self._scene_ref = scene
            if previous is not None:
                previous.close()
        return self._device_scene

    def replace_scene(self, scene):
        """Make ``scene`` resident; the previous one stays if the upload fails."""
        with self._lock:
            if self._closed:
                raise RuntimeError("renderer is closed")
            self._resident(scene)

    def close(self):
        with self._lock:
            if self._closed:
                return
            self._closed = True
            if self._device_scene is not None:
                self._device_scene.close()
                self._device_scene = None
            if self._context is not None:
                self._context.close()
                self._context = None


def capture_motion_state(scene, camera, size):
    """What the next frame's motion output needs of this one: the scene
    object, camera, image size and each visible mesh's vertices, indices
    and transform (copies), keyed by mesh id."""
    return {"scene": scene, "camera": camera, "size": tuple(size),
            "meshes": {mesh.id: (np.array(mesh.vertices, copy=True), np.array(mesh.indices, copy=True),
                                 np.array(mesh.transform.matrix, copy=True)) for mesh in scene.visible_meshes}}


def project_positions(positions, camera, width, height):
    """Pixel coordinates of world points under a look-at camera with a
    vertical field of view, and whether each lies in front of it (float64,
    as the reference's host projection, targets/vulkan/api.py:1118-1137)."""
    origin = np.asarray(camera.position, np.float64)
    forward = np.asarray(camera.target, np.float64) - origin
    forward = forward / np.linalg.norm(forward)
    right = np.cross(forward, np.asarray(camera.up, np.float64))
    right = right / np.linalg.norm(right)
    up = np.cross(right, forward)
    rel = np.asarray(positions, np.float64) - origin
    depth = rel @ forward
    half = math.tan(math.radians(float(camera.vertical_fov_degrees)) * 0.5)
    z = np.maximum(depth, 1e-12)
    x = (rel @ right) / (z * half * (width / height))
    y = -(rel @ up) / (z * half)
    return np.stack(((x + 1.0) * 0.5 * width, (y + 1.0) * 0.5 * height), axis=1), depth > 1e-8


def motion_vectors(scene, camera, width, height, primitive, position, barycentric, previous):
    """Per-pixel motion of sample 0's primary hit since the previous frame
    (the reference's opt-in ``motion`` product, targets/vulkan/api.py:1151-1200):
    the hit's barycentric point on its triangle as the previous frame's mesh
    vertices and transform placed it, projected by the previous camera, is
    subtracted from the hit's projection now.  Zero without a previous frame
    of the same scene object and size, on misses, for meshes whose triangles
    changed, and where either projection is behind its camera."""
    motion = np.zeros((height, width, 2), np.float32)
    if previous is None or previous["scene"] is not scene or previous["size"] != (width, height):
        return motion
    prim = np.asarray(primitive).reshape(-1)
    pos = np.asarray(position).reshape(-1, 3)
    bary = np.asarray(barycentric).reshape(-1, 2)
    flat = motion.reshape(-1, 2)
    first = 0
    for mesh in scene.visible_meshes:
        count = len(mesh.indices)
        rows = np.flatnonzero((prim >= first) & (prim < first + count))
        old = previous["meshes"].get(mesh.id)
        if rows.size and old is not None:
            vertices, indices, transform = old
            if indices.shape == np.shape(mesh.indices) and np.array_equal(indices, mesh.indices) and \
                    vertices.shape == np.shape(mesh.vertices):
                corners = vertices[indices[prim[rows].astype(np.int64) - first]]
                u, v = bary[rows, 0], bary[rows, 1]
                weights = np.stack((1.0 - u - v, u, v), axis=1)
                local = np.einsum("nk,nkc->nc", weights, corners)
                world = local @ transform[:3, :3].T + transform[:3, 3]
                now, ok_now = project_positions(pos[rows], camera, width, height)
                then, ok_then = project_positions(world, previous["camera"], width, height)
                ok = ok_now & ok_then
                flat[rows[ok]] = (now[ok] - then[ok]).astype(np.float32)
        first += count
    return motion


def _check_size(width, height):
    width, height = int(width), int(height)
    if width < 1 or height < 1:
        raise ValueError("width and height must be positive")
    return width, height


class HipPathTracer(_ResidentMixin, RendererImplementation):
    """GI wavefront path tracer on one MI355X (drop-in for the Vulkan GI path)."""

    available_outputs = ("color", "variance", "depth", "normal", "instance_id", "object_id", "material_id", "motion")
    implementation = RendererImplementationInfo(
        name="hip-wavefront-gi", family="global_illumination", graphics_api="hip",
    )

    def __init__(self, config=None, **options):
        if config is not None and options:
            raise TypeError("pass config or renderer options, not both")
        self.config = config or HipRendererConfig(**options)
        if not isinstance(self.config, HipRendererConfig):
            raise TypeError("config must be a HipRendererConfig")
        self._init_resident(self.config.device_index, self.config.bvh_leaf_size, self.config.bvh_builder)
        self.device = f"HIP device {self.config.device_index}"
        self.last_timings = {}
        self.last_stats = None
        self._output_history = None

    def reset_output_history(self):
        """Discard the previous frame's state that the motion output uses."""
        self._output_history = None

    def replace_scene(self, scene):
        _ResidentMixin.replace_scene(self, scene)
        self.reset_output_history()

    @property
    def capabilities(self):
        return {
            "renderer": self.implementation.name,
            "outputs": self.available_outputs,
            "features": frozenset({"offscreen_rendering", "global_illumination", "instancing",
                                   "area_lights", "transmission", "resident_scene", "picking", "restir_di"}),
            "limits": {"max_bounces": 64, "max_samples_per_pixel": int(self.config.max_samples_per_call)},
            "device": self.device,
        }

    def reconfigure(self, **changes):
        """Apply hot settings in place; structural changes raise RuntimeError."""
        unknown = set(changes) - {f.name for f in fields(HipRendererConfig)}
        if unknown:
            raise ValueError(f"unknown settings: {sorted(unknown)}")
        structural = set(changes) - HOT_SETTINGS
        if structural:
            raise RuntimeError(f"settings {sorted(structural)} require renderer recreation")
        with self._lock:
            self.config = replace(self.config, **changes)
            return self.config

    def _render(self, scene, camera, width, height, samples, frame_index, keep_on_device=False, products=None):
        if self._closed:
            raise RuntimeError("renderer is closed")
        width, height = _check_size(width, height)
        samples = self.config.samples_per_pixel if samples is None else int(samples)
        if not 1 <= samples <= int(self.config.max_samples_per_call):
            raise ValueError(f"samples must be between 1 and {self.config.max_samples_per_call}")
        frame_index = int(frame_index)
        if frame_index < 0:
            raise ValueError("frame_index cannot be negative")
        started = time.perf_counter()
        device_scene = self._resident(scene)
        context = self._ensure_context()
        lib = N.library()
        N.check(lib.olhip_set_profiling(context.handle, int(bool(self.config.profiling))))
        N.check(lib.olhip_set_traversal(context.handle, int(self.config.traversal_lds_stack),
                                        int(self.config.traversal_block)))
        N.check(lib.olhip_set_work_counters(context.handle, int(bool(self.config.work_counters))))
        N.check(lib.olhip_set_fused_primary(context.handle, int(self.config.fused_primary)))
        position, forward, right, up = camera_basis(camera)
        cam = N.Camera()
        cam.position[:] = position.tolist()
        cam.forward[:] = forward.tolist()
        cam.right[:] = right.tolist()
        cam.up[:] = up.tolist()
        cam.projection = camera_projection(camera)
        out = None if keep_on_device else np.empty((height, width, 4), np.float32)
        stats = N.Stats()
        settings = self.config.gi_settings()
        if products is None:
            N.check(lib.olhip_render_gi(context.handle, device_scene.handle, N.ctypes.byref(cam), width, height,
                                        samples, frame_index, N.ctypes.byref(settings),
                                        N.ptr(out) if out is not None else None, N.ctypes.byref(stats)))
        else:
            N.check(lib.olhip_render_products(
                context.handle, device_scene.handle, N.ctypes.byref(cam), width, height, samples, frame_index,
                N.ctypes.byref(settings), N.ptr(out) if out is not None else None,
                *(N.ptr(products[k], N.ctypes.c_uint32 if k == "primitive" else N.ctypes.c_float)
                  if k in products else None for k in ("depth", "normal", "primitive", "variance")),
                N.ctypes.byref(stats)))
        if stats.queue_overflow:
            # The engine fails such a job itself; a render never returns an
            # image from which paths or shadow rays were dropped.
            raise RuntimeError(f"ray queue overflow: {stats.queue_overflow} continuations or shadow groups dropped")
        total_ms = (time.perf_counter() - started) * 1000.0
        rays = stats.primary_rays + stats.extension_rays + stats.shadow_rays
        self.last_stats = stats
        self.last_timings = {
            "total_ms": total_ms, "gpu_ms": stats.total_ms,
            "trace_ms": stats.trace_ms, "shade_ms": stats.shade_ms, "shadow_ms": stats.shadow_ms,
            "primary_rays": int(stats.primary_rays), "extension_rays": int(stats.extension_rays),
            "shadow_rays": int(stats.shadow_rays),
            "mrays_per_second": rays / (stats.total_ms * 1e3) if stats.total_ms > 0 else 0.0,
            **device_scene.timings,
        }
        return out

    def render_frame(self, scene, camera, width, height, *, samples=None, frame_index=0):
        with self._lock:
            return self._render(scene, camera, width, height, samples, frame_index)

    render_wavefront = render_frame

    def render_products(self, scene, camera, width, height, *, outputs=("color",), samples=None, frame_index=0):
        outputs = tuple(outputs)
        unsupported = set(outputs) - set(self.available_outputs)
        if not outputs or unsupported:
            raise ValueError(f"unsupported render outputs: {tuple(sorted(unsupported))}")
        if len(set(outputs)) != len(outputs):
            raise ValueError(f"render outputs listed twice: {outputs}")
        width, height = _check_size(width, height)
        ids = {"instance_id", "object_id", "material_id"} & set(outputs)
        buffers = {}
        if {"depth"} & set(outputs):
            buffers["depth"] = np.empty((height, width), np.float32)
        if "normal" in outputs:
            buffers["normal"] = np.empty((height, width, 3), np.float32)
        if ids:
            buffers["primitive"] = np.empty((height, width), np.uint32)
        if "variance" in outputs:
            buffers["variance"] = np.empty((height, width), np.float32)
        with self._lock:
            color = self._render(scene, camera, width, height, samples, frame_index, products=buffers)
            if "motion" in outputs:
                # Sample 0's primary hits with their barycentrics: the render's
                # own camera rays against the resident hierarchy (the same
                # closest-hit rule, so the same hits as the primitive product).
                o, d = self._ensure_context().camera_rays(camera, width, height, int(frame_index), 0)
                t, u, v, prim = self._resident(scene).trace(o[:, :3], d[:, :3], float(o[0, 3]), float(d[0, 3]))
                position = o[:, :3] + t[:, None] * d[:, :3]
                motion = motion_vectors(scene, camera, width, height, prim, position, np.stack((u, v), axis=1),
                                        self._output_history)
                self._output_history = capture_motion_state(scene, camera, (width, height))
        result = {}
        if "motion" in outputs:
            result["motion"] = motion
        if "color" in outputs:
            result["color"] = color
        for name in ("depth", "normal", "variance"):
            if name in outputs:
                result[name] = buffers[name]
        if ids:
            # Primitive -> id maps on the host, as the reference does
            # (targets/vulkan/api.py:1360-1386); misses stay 0xFFFFFFFF.
            prim = buffers["primitive"]
            hit = prim != np.uint32(0xFFFFFFFF)
            maps = {"instance_id": scene.triangle_instance_ids, "object_id": scene.triangle_instance_ids,
                    "material_id": scene.triangle_material_ids}
            for name in ids:
                ids_out = np.full((height, width), np.uint32(0xFFFFFFFF), np.uint32)
                table = np.asarray(maps[name](), np.uint32)
                ids_out[hit] = table[prim[hit]]
                result[name] = ids_out
        return {name: result[name] for name in outputs}

    def render_output(self, scene, camera, width, height, *, pixel_format="rgba8", samples=None, frame_index=0,
                      output_size=None, device_out=None):
        """Display frame of a render (the reference's GPU-resident output,
        ``targets/vulkan/api.py:1074-1115``): tone-mapped RGBA8
        ``uint8 [H, W, 4]``, NV12 ``uint8 [H * 3 / 2, W]`` or P010
        ``uint16 [H * 3 / 2, W]``.  ``output_size`` scales P010 only.  With
        ``device_out`` (an address of device memory on the renderer's device,
        e.g. a torch tensor's ``data_ptr()``), the frame is written there and
        nothing returns to the host."""
        pixel_format = str(pixel_format).lower()
        if pixel_format not in OUTPUT_FORMATS:
            raise ValueError("pixel_format must be 'rgba8', 'nv12', or 'p010'")
        width, height = _check_size(width, height)
        ow, oh = _check_size(*(output_size or (width, height)))
        if pixel_format != "p010" and (ow, oh) != (width, height):
            raise ValueError("only P010 output can be scaled")
        if pixel_format != "rgba8" and (ow % 4 or oh % 2):
            raise ValueError(f"{pixel_format.upper()} output requires width divisible by 4 and even height")
        with self._lock:
            self._render(scene, camera, width, height, samples, frame_index, keep_on_device=True)
            lib = N.library()
            fmt = OUTPUT_FORMATS[pixel_format]
            if device_out is not None:
                N.check(lib.olhip_frame_output(self._ensure_context().handle, fmt, float(self.config.wavefront_exposure),
                                               ow, oh, N.ctypes.c_void_p(int(device_out)), 1))
                return None
            if pixel_format == "rgba8":
                out = np.empty((oh, ow, 4), np.uint8)
            else:
                out = np.empty((oh * 3 // 2, ow), np.uint8 if pixel_format == "nv12" else np.uint16)
            N.check(lib.olhip_frame_output(self._ensure_context().handle, fmt, float(self.config.wavefront_exposure),
                                           ow, oh, out.ctypes.data_as(N.ctypes.c_void_p), 0))
            return out

    def pick(self, scene, camera, viewport_size, pixel, *, options=None, mapping=None):
        """Object under a pixel: the pixel's primary ray of frame 0, sample 0
        (the jittered ray a render traces) against the resident BVH, as
        ``targets/vulkan/api.py:936-987``.  None on a miss."""
        options = PickOptions() if options is None else options
        if not isinstance(options, PickOptions):
            raise TypeError("options must be PickOptions")
        if options != PickOptions():
            raise NotImplementedError("pick policies other than the default are not supported on the HIP renderer")
        if mapping is not None:
            if not isinstance(mapping, ViewportMapping):
                raise TypeError("mapping must be ViewportMapping")
            pixel = mapping.map_pixel(pixel)
            if pixel is None:
                return None
            viewport_size = mapping.render_size
        width, height = _check_size(*viewport_size)
        px, py = (float(v) for v in pixel)
        if not (math.isfinite(px) and math.isfinite(py)):
            raise ValueError("pixel must contain two finite coordinates")
        x = min(max(int(math.floor(px)), 0), width - 1)
        y = min(max(int(math.floor(py)), 0), height - 1)
        with self._lock:
            if self._closed:
                raise RuntimeError("renderer is closed")
            device_scene = self._resident(scene)
            o, d = self._ensure_context().camera_rays(camera, width, height, 0, 0, (x, y, 1, 1))
            t, u, v, prim = device_scene.trace(o[:, :3], d[:, :3], float(o[0, 3]), float(d[0, 3]))
        primitive = int(prim[0])
        if primitive == 0xFFFFFFFF:
            return None
        object_id = int(np.asarray(scene.triangle_instance_ids())[primitive])
        start, _ = scene.object_triangle_range(object_id)
        distance = np.float32(t[0])
        position = o[0, :3] + distance * d[0, :3]
        return PickResult(scene.get_instance(object_id), object_id, float(distance),
                          tuple(float(c) for c in position), primitive, primitive - int(start),
                          (float(np.float32(1.0) - u[0] - v[0]), float(u[0]), float(v[0])))

    def close(self):
        _ResidentMixin.close(self)


# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class HipReferenceConfig:
    """``ReferenceConfig`` (``renderers/reference/cpu.py:9-21``) + device."""

    samples_per_pixel: int = 1
    max_bounces: int = 3
    seed: int = 1
    device_index: int = 0
    bvh_leaf_size: int = 2
    # Index-order triangle scan below this triangle count (exact tie order).
    brute_force_below: int = 257

    def __post_init__(self):
        if not 1 <= int(self.samples_per_pixel) <= 4096:
            raise ValueError("samples_per_pixel must be between 1 and 4096")
        if not 1 <= int(self.max_bounces) <= 64:
            raise ValueError("max_bounces must be between 1 and 64")
        if int(self.seed) < 0:
            raise ValueError("seed cannot be negative")


def reference_camera(camera, width, height):
    """float32 basis and float64 projection scalars as ``_camera_rays`` computes them."""
    position = np.asarray(camera.position, dtype=np.float32)
    forward = np.asarray(camera.target, dtype=np.float32) - position
    forward /= np.linalg.norm(forward)
    right = np.cross(forward, np.asarray(camera.up, dtype=np.float32))
    right /= np.linalg.norm(right)
    up = np.cross(right, forward)
    aspect = width / height
    projection = camera_projection(camera)
    if projection == 2:
        k_x = math.radians(camera.horizontal_fov_degrees)
        k_y = math.radians(camera.vertical_fov_degrees)
    elif projection == 1:
        half = camera.vertical_size * 0.5
        k_x, k_y = aspect * half, half
    else:
        half = math.tan(math.radians(camera.vertical_fov_degrees) / 2.0)
        k_x, k_y = aspect * half, half
    cam = N.E0Camera()
    cam.position[:] = position.tolist()
    cam.forward[:] = forward.tolist()
    cam.right[:] = right.tolist()
    cam.up[:] = up.tolist()
    cam.k_x, cam.k_y = float(k_x), float(k_y)
    cam.projection = projection
    return cam


class HipReferenceRenderer(_ResidentMixin, RendererImplementation):
    """The reference's diffuse CPU estimator, run on the GPU.

    The PCG64 stream is drawn on the host in the reference's order; rays,
    intersection, sky, compaction and shading run in ``libolhip.so``.
    """

    available_outputs = ("color",)
    implementation = RendererImplementationInfo(
        name="hip-reference", family="reference", graphics_api="hip",
    )

    def __init__(self, config=None, **options):
        if config is not None and options:
            raise TypeError("pass config or reference options, not both")
        self.config = config or HipReferenceConfig(**options)
        if not isinstance(self.config, HipReferenceConfig):
            raise TypeError("config must be a HipReferenceConfig")
        self._init_resident(self.config.device_index, self.config.bvh_leaf_size)
        self.device = f"HIP device {self.config.device_index}"
        self.last_timings = {}
        self.last_primary_hits = None

    @property
    def capabilities(self):
        return {
            "renderer": self.implementation.name,
            "outputs": self.available_outputs,
            "features": frozenset({"offscreen_rendering", "instancing"}),
            "limits": {"max_bounces": 64, "max_samples_per_pixel": 4096},
            "device": self.device,
        }

    def render_frame(self, scene, camera, width, height, *, samples=None, frame_index=0,
                     capture_primary_hits=False):
        with self._lock:
            if self._closed:
                raise RuntimeError("reference renderer is closed")
            samples = self.config.samples_per_pixel if samples is None else int(samples)
            width, height = int(width), int(height)
            if width <= 0 or height <= 0:
                raise ValueError("width and height must be positive")
            if samples <= 0:
                raise ValueError("samples and max_bounces must be positive")
            if getattr(scene, "visible_volumes", ()):
                raise RuntimeError("volumes are not supported by the HIP reference renderer")
            started = time.perf_counter()
            device_scene = self._resident(scene)
            context = self._ensure_context()
            lib = N.library()
            cam = reference_camera(camera, width, height)
            env = N.E0Env()
            environment = getattr(scene, "environment", None)
            image = None
            if environment is None:
                env.mode = 0
            else:
                env.color[:] = np.asarray(environment.color, np.float32).tolist()
                env.intensity = float(np.float32(float(environment.intensity)))
                env.rotation = float(environment.rotation)
                if environment.image is None:
                    env.mode = 1
                else:
                    env.mode = 2
                    image = np.ascontiguousarray(environment.image[..., :3], np.float32)
                    env.image = N.ptr(image)
                    env.height, env.width = image.shape[:2]
            brute = int(device_scene.triangle_count < int(self.config.brute_force_below))
            N.check(lib.olhip_e0_begin(context.handle, device_scene.handle, N.ctypes.byref(cam), width, height,
                                       N.ctypes.byref(env), brute))
            rng = np.random.default_rng(int(self.config.seed) + int(frame_index))
            n_hit = N.ctypes.c_uint32(0)
            n_active = N.ctypes.c_uint32(0)
            primary = None
            rays = 0
            for sample in range(samples):
                jx = rng.random((height, width), dtype=np.float32)
                jy = rng.random((height, width), dtype=np.float32)
                N.check(lib.olhip_e0_sample(context.handle, N.ptr(jx), N.ptr(jy)))
                for bounce in range(int(self.config.max_bounces)):
                    N.check(lib.olhip_e0_trace(context.handle, N.ctypes.byref(n_hit), N.ctypes.byref(n_active)))
                    if n_active.value == 0:
                        break
                    rays += n_active.value
                    if capture_primary_hits and sample == 0 and bounce == 0:
                        primary = np.empty(n_active.value, np.int32)
                        N.check(lib.olhip_e0_last_hits(context.handle, N.ptr(primary, N.ctypes.c_int32),
                                                       n_active.value))
                    count = n_hit.value
                    if count:
                        deviates = rng.normal(size=(count, 3)).astype(np.float32)
                        N.check(lib.olhip_e0_shade(context.handle, N.ptr(deviates), count))
                    else:
                        N.check(lib.olhip_e0_shade(context.handle, None, 0))
                N.check(lib.olhip_e0_sample_end(context.handle))
            out = np.empty((height, width, 4), np.float32)
            N.check(lib.olhip_e0_end(context.handle, samples, N.ptr(out)))
            self.last_primary_hits = primary
            self.last_timings = {"total_ms": (time.perf_counter() - started) * 1000.0, "rays": rays,
                                 **device_scene.timings}
            return out

    render_wavefront = render_frame

    def close(self):
        _ResidentMixin.close(self)


__all__ = [
    "DeviceContext", "DeviceScene", "HOT_SETTINGS", "HipPathTracer", "HipReferenceConfig",
    "HipReferenceRenderer", "HipRendererConfig", "reference_camera",
]
