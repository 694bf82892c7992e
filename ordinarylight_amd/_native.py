"""ctypes binding of ``libolhip.so`` (declarations in ``include/olhip.h``).

The library is the only compute path of this package: importing this module
never falls back to anything else, and :func:`library` raises when the
shared object is missing or does not load.
"""

from __future__ import annotations

import ctypes
import os
import pathlib
import threading

HERE = pathlib.Path(__file__).resolve().parent
LIBRARY_PATH = pathlib.Path(os.environ.get("OLHIP_LIBRARY", HERE / "libolhip.so"))

OLHIP_OK = 0
OLHIP_ERR_VALUE = -1
OLHIP_ERR_RUNTIME = -2
OLHIP_ERR_MEMORY = -3
OLHIP_ERR_DEVICE = -4
OLHIP_TRACE_SCRATCH_BYTES = 16 << 20

c_u32 = ctypes.c_uint32
c_f32 = ctypes.c_float
c_f64 = ctypes.c_double
P_f32 = ctypes.POINTER(ctypes.c_float)
P_u32 = ctypes.POINTER(ctypes.c_uint32)
P_i32 = ctypes.POINTER(ctypes.c_int32)


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("vertex_count", c_u32), ("positions", P_f32), ("normals", P_f32),
        ("triangle_count", c_u32), ("indices", P_u32), ("material_ids", P_u32),
        ("material_count", c_u32), ("materials", P_f32),
        ("analytic_light_count", c_u32), ("analytic_lights", P_f32),
        ("area_light_count", c_u32), ("area_lights", P_f32),
        ("area_light_weight", c_f32), ("leaf_size", c_u32), ("builder", c_u32),
        ("texture_word_count", c_u32), ("texture_words", P_u32), ("binding_count", c_u32),
        ("bindings", P_f32), ("uvs", P_f32), ("tangents", P_f32), ("uv_density", P_f32),
        ("face_params", P_f32),
    ]


BUILDER_LBVH, BUILDER_SAH = 0, 1


class InstancedDesc(ctypes.Structure):
    _fields_ = [
        ("resource_count", c_u32), ("vertex_offsets", P_u32), ("triangle_offsets", P_u32),
        ("positions", P_f32), ("normals", P_f32), ("indices", P_u32),
        ("instance_count", c_u32), ("instance_resource", P_u32), ("instance_material", P_u32),
        ("instance_triangle_offset", P_u32), ("instance_id", P_u32), ("world_to_object", P_f32),
        ("normal_matrix", P_f32), ("determinant", P_f32), ("world_bounds", P_f32),
        ("material_count", c_u32), ("materials", P_f32),
        ("analytic_light_count", c_u32), ("analytic_lights", P_f32),
        ("area_light_count", c_u32), ("area_lights", P_f32),
        ("area_light_weight", c_f32), ("leaf_size", c_u32), ("builder", c_u32),
        ("world_spheres", P_f32), ("texture_word_count", c_u32), ("texture_words", P_u32),
    ]


class SceneInfo(ctypes.Structure):
    _fields_ = [
        ("triangle_count", c_u32), ("node_count", c_u32), ("bvh_depth", c_u32),
        ("stack_depth", c_u32), ("build_ms", c_f64), ("device_bytes", ctypes.c_uint64),
    ]


class Camera(ctypes.Structure):
    _fields_ = [
        ("position", c_f32 * 3), ("forward", c_f32 * 3), ("right", c_f32 * 3),
        ("up", c_f32 * 3), ("projection", c_u32),
    ]


class GiSettings(ctypes.Structure):
    _fields_ = [
        ("max_bounces", c_u32), ("secondary_area_samples", c_u32),
        ("environment_samples", c_u32), ("russian_roulette_start", c_u32),
        ("russian_roulette_min_survival", c_f32), ("secondary_nee_probability", c_f32),
        ("unified_secondary_nee", c_u32), ("path_capacity", c_u32),
        ("restir_candidates", c_u32), ("restir_spatial_neighbors", c_u32), ("restir_spatial_radius", c_u32),
        ("restir_history_limit", c_u32), ("restir_pairwise_mis", c_u32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("primary_rays", ctypes.c_uint64), ("extension_rays", ctypes.c_uint64),
        ("shadow_rays", ctypes.c_uint64), ("queue_overflow", ctypes.c_uint64),
        ("total_ms", c_f64), ("trace_ms", c_f64), ("shade_ms", c_f64), ("shadow_ms", c_f64),
        ("trace_launches", c_u32), ("shade_launches", c_u32), ("shadow_launches", c_u32),
        ("chunks", c_u32), ("closest_nodes", ctypes.c_uint64), ("closest_tris", ctypes.c_uint64),
        ("shadow_nodes", ctypes.c_uint64), ("shadow_tris", ctypes.c_uint64),
        ("closest_instances", ctypes.c_uint64), ("shadow_instances", ctypes.c_uint64),
        ("shaded_hits", ctypes.c_uint64), ("area_samples", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class EngineConfig(ctypes.Structure):
    _fields_ = [
        ("path_slots", ctypes.c_uint64), ("entry_factor", c_u32), ("lds_stack", ctypes.c_int32),
        ("trace_block", ctypes.c_int32), ("use_graph", ctypes.c_int32), ("append_limit", c_u32),
        ("fused_primary", c_u32),
    ]


class JobDesc(ctypes.Structure):
    _fields_ = [
        ("scene", ctypes.c_void_p), ("camera", Camera), ("width", c_u32), ("height", c_u32),
        ("samples", c_u32), ("frame_index", c_u32), ("settings", GiSettings), ("products", c_u32), ("pad", c_u32),
    ]


class JobStats(ctypes.Structure):
    _fields_ = [
        ("primary_rays", ctypes.c_uint64), ("extension_rays", ctypes.c_uint64), ("shadow_rays", ctypes.c_uint64),
        ("steps", c_u32), ("chunks", c_u32), ("submit_ms", c_f64), ("admit_ms", c_f64), ("done_ms", c_f64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class EngineStats(ctypes.Structure):
    _fields_ = [
        ("steps", ctypes.c_uint64), ("jobs_done", ctypes.c_uint64), ("rays", ctypes.c_uint64),
        ("queue_overflow", ctypes.c_uint64), ("busy_ms", c_f64), ("resident_jobs", c_u32),
        ("queued_jobs", c_u32), ("max_resident_jobs", c_u32), ("fused_steps", c_u32),
        ("path_slots", ctypes.c_uint64), ("entry_slots", ctypes.c_uint64), ("device_ms", c_f64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad"}


JOB_UNKNOWN, JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED = -1, 0, 1, 2, 3
PRODUCT_PRIMARY, PRODUCT_VARIANCE = 1, 2


class E0Camera(ctypes.Structure):
    _fields_ = [
        ("position", c_f32 * 3), ("forward", c_f32 * 3), ("right", c_f32 * 3),
        ("up", c_f32 * 3), ("k_x", c_f64), ("k_y", c_f64), ("projection", c_u32),
    ]


class E0Env(ctypes.Structure):
    _fields_ = [
        ("mode", c_u32), ("color", c_f32 * 3), ("intensity", c_f32), ("rotation", c_f64),
        ("image", P_f32), ("width", c_u32), ("height", c_u32),
    ]


_SIGNATURES = {
    "olhip_last_error": (ctypes.c_char_p, []),
    "olhip_version": (ctypes.c_int, []),
    "olhip_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "olhip_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "olhip_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "olhip_scene_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SceneDesc), ctypes.POINTER(ctypes.c_void_p)]),
    "olhip_scene_create_instanced": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(InstancedDesc),
                                                    ctypes.POINTER(ctypes.c_void_p)]),
    "olhip_scene_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "olhip_scene_info_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SceneInfo)]),
    "olhip_scene_bvh": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "olhip_render_gi": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Camera), c_u32, c_u32,
                                       c_u32, c_u32, ctypes.POINTER(GiSettings), P_f32, ctypes.POINTER(Stats)]),
    "olhip_frame_download": (ctypes.c_int, [ctypes.c_void_p, P_f32]),
    "olhip_bandwidth_probe": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, c_u32, ctypes.POINTER(c_f64)]),
    "olhip_set_profiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "olhip_set_traversal": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "olhip_set_fused_primary": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "olhip_engine_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(EngineConfig),
                                           ctypes.POINTER(ctypes.c_void_p)]),
    "olhip_engine_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "olhip_engine_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(JobDesc),
                                           ctypes.POINTER(ctypes.c_uint64)]),
    "olhip_engine_step": (ctypes.c_int, [ctypes.c_void_p, P_u32]),
    "olhip_engine_run": (ctypes.c_int, [ctypes.c_void_p]),
    "olhip_engine_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]),
    "olhip_engine_wait": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, c_f64, ctypes.POINTER(ctypes.c_int)]),
    "olhip_engine_result": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, P_f32, ctypes.POINTER(JobStats)]),
    "olhip_engine_products": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, P_f32, P_f32, P_u32, P_f32]),
    "olhip_render_products": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Camera), c_u32, c_u32,
                                             c_u32, c_u32, ctypes.POINTER(GiSettings), P_f32, P_f32, P_f32, P_u32,
                                             P_f32, ctypes.POINTER(Stats)]),
    "olhip_engine_start": (ctypes.c_int, [ctypes.c_void_p]),
    "olhip_engine_stop": (ctypes.c_int, [ctypes.c_void_p]),
    "olhip_engine_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(EngineStats)]),
    "olhip_set_work_counters": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "olhip_trace": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, P_f32, P_f32, c_u32, P_f32]),
    "olhip_trace_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, c_u32,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "olhip_frame_output": (ctypes.c_int, [ctypes.c_void_p, c_u32, ctypes.c_float, c_u32, c_u32, ctypes.c_void_p,
                                          c_u32]),
    "olhip_scene_update": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(SceneDesc)]),
    "olhip_scene_update_instanced": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(InstancedDesc)]),
    "olhip_sah_topology": (ctypes.c_int, [P_f32, c_u32, ctypes.POINTER(c_u32), ctypes.POINTER(c_u32),
                                          ctypes.POINTER(c_u32), ctypes.POINTER(c_u32), ctypes.POINTER(c_u32)]),
    "olhip_camera_rays": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Camera), c_u32, c_u32, c_u32, c_u32, c_u32,
                                         c_u32, c_u32, c_u32, P_f32, P_f32]),
    "olhip_e0_begin": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(E0Camera), c_u32, c_u32,
                                      ctypes.POINTER(E0Env), c_u32]),
    "olhip_e0_sample": (ctypes.c_int, [ctypes.c_void_p, P_f32, P_f32]),
    "olhip_e0_trace": (ctypes.c_int, [ctypes.c_void_p, P_u32, P_u32]),
    "olhip_e0_shade": (ctypes.c_int, [ctypes.c_void_p, P_f32, c_u32]),
    "olhip_e0_sample_end": (ctypes.c_int, [ctypes.c_void_p]),
    "olhip_e0_end": (ctypes.c_int, [ctypes.c_void_p, c_u32, P_f32]),
    "olhip_e0_last_hits": (ctypes.c_int, [ctypes.c_void_p, P_i32, c_u32]),
}

EXPORTED_SYMBOLS = tuple(_SIGNATURES)

_lock = threading.Lock()
_lib = None


def library():
    """Load ``libolhip.so`` once; raises ``RuntimeError`` if it is unavailable."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIBRARY_PATH.exists():
                raise RuntimeError(
                    f"the HIP renderer library is missing ({LIBRARY_PATH}); "
                    "build it with `python -m ordinarylight_amd.build_native`"
                )
            lib = ctypes.CDLL(str(LIBRARY_PATH))
            for name, (restype, argtypes) in _SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = restype
                fn.argtypes = argtypes
            _lib = lib
        return _lib


def check(status):
    """Raise the exception class the status mirrors."""
    if status == OLHIP_OK:
        return
    message = library().olhip_last_error().decode("utf-8", "replace") or f"olhip status {status}"
    if status == OLHIP_ERR_VALUE:
        raise ValueError(message)
    if status == OLHIP_ERR_MEMORY:
        raise MemoryError(message)
    raise RuntimeError(message)


def ptr(array, ctype=ctypes.c_float):
    """ctypes pointer to a C-contiguous NumPy array (or None)."""
    if array is None:
        return None
    if not array.flags.c_contiguous:
        raise ValueError("array must be C-contiguous")
    return array.ctypes.data_as(ctypes.POINTER(ctype))


def device_count():
    count = ctypes.c_int(0)
    check(library().olhip_device_count(ctypes.byref(count)))
    return count.value
