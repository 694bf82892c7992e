"""Batched and served rendering on one MI355X.

:class:`BatchEngine` drives the library's render engine (``olhip_engine_*``):
many jobs of different scenes, image sizes, sample counts and bounce limits
share one pool of resident paths and one set of ray/shadow queues, and jobs
join while others are in flight (continuous batching).  Each job's image is
bit-identical to rendering it alone: every path depends only on (pixel,
sample, frame) and the float64 pixel sums run in sample order.

:class:`RenderServer` keeps scenes resident, runs the engine's worker thread
and lets callers submit jobs and wait for results (p50/p99 latency from the
engine's own clock).  :func:`autotune` picks the path-slot capacity and the
traversal launch shape by timing a calibration batch.

The reference has no batching or serving layer: its executor renders one
frame at a time (targets/vulkan/core.py:2338-3215); the per-job semantics are
those of ``render_frame`` (targets/vulkan/api.py:1233-1451).
"""

from __future__ import annotations

import dataclasses
import itertools
import math
import threading
import time

import numpy as np

from . import _native as N
from .cameras import camera_basis, camera_projection
from .renderers.hip import DeviceContext, DeviceScene, HipRendererConfig

__all__ = [
    "EngineSettings", "Job", "BatchEngine", "RenderServer", "plain_render", "autotune",
    "latency_percentiles", "poisson_arrivals",
]


@dataclasses.dataclass(frozen=True)
class EngineSettings:
    """Engine capacity and launch shape (0 = library default)."""

    path_slots: int = 0          # resident paths shared by all jobs (default 2^25)
    entry_factor: int = 0        # shadow-entry slots per path slot (default 2)
    lds_stack: int = 16          # traversal stack entries kept in LDS
    trace_block: int = 128       # traversal workgroup size
    use_graph: bool = True       # replay each step as a captured hipGraph
    # Failure-path testing only: continuations a queue segment accepts per
    # step (0 = its full size, which cannot overflow).
    append_limit: int = 0
    # A step of fresh camera rays only (flat, untextured scenes without
    # ReSTIR) traces and shades them in one kernel (the hybrid strategy).
    fused_primary: bool = False

    def native(self):
        c = N.EngineConfig()
        c.path_slots = int(self.path_slots)
        c.entry_factor = int(self.entry_factor)
        c.lds_stack = int(self.lds_stack)
        c.trace_block = int(self.trace_block)
        c.use_graph = int(bool(self.use_graph))
        c.append_limit = int(self.append_limit)
        c.fused_primary = int(bool(self.fused_primary))
        return c


@dataclasses.dataclass
class Job:
    """One render request."""

    scene: DeviceScene
    camera: object
    width: int
    height: int
    samples: int = 1
    frame_index: int = 0
    config: HipRendererConfig | None = None   # estimator settings (max_bounces, NEE, ...)
    max_bounces: int | None = None            # shorthand overriding config.max_bounces
    products: int = 0                         # N.PRODUCT_PRIMARY | N.PRODUCT_VARIANCE

    def settings(self):
        cfg = self.config or HipRendererConfig(max_bounces=self.max_bounces or 5,
                                               samples_per_pixel=min(max(int(self.samples), 1), 64),
                                               max_samples_per_call=max(64, int(self.samples)))
        if self.max_bounces is not None and cfg.max_bounces != self.max_bounces:
            cfg = dataclasses.replace(cfg, max_bounces=int(self.max_bounces))
        s = cfg.gi_settings()
        s.path_capacity = 0
        return s

    def native(self):
        if not 1 <= int(self.samples) <= 65536:
            raise ValueError("samples must be between 1 and 65536")
        d = N.JobDesc()
        d.scene = self.scene.handle
        position, forward, right, up = camera_basis(self.camera)
        d.camera.position[:] = position.tolist()
        d.camera.forward[:] = forward.tolist()
        d.camera.right[:] = right.tolist()
        d.camera.up[:] = up.tolist()
        d.camera.projection = camera_projection(self.camera)
        d.width = int(self.width)
        d.height = int(self.height)
        d.samples = int(self.samples)
        d.frame_index = int(self.frame_index)
        d.settings = self.settings()
        d.products = int(self.products)
        return d


class BatchEngine:
    """One ``olhip_engine`` on a device context."""

    def __init__(self, context: DeviceContext, settings: EngineSettings | None = None):
        self._lib = N.library()
        self.context = context
        self.settings = settings or EngineSettings()
        handle = N.ctypes.c_void_p()
        cfg = self.settings.native()
        N.check(self._lib.olhip_engine_create(context.handle, N.ctypes.byref(cfg), N.ctypes.byref(handle)))
        self.handle = handle
        self._jobs = {}
        self._worker = False

    def submit(self, job: Job) -> int:
        desc = job.native()
        jid = N.ctypes.c_uint64()
        N.check(self._lib.olhip_engine_submit(self.handle, N.ctypes.byref(desc), N.ctypes.byref(jid)))
        self._jobs[int(jid.value)] = job
        return int(jid.value)

    def step(self) -> int:
        pending = N.c_u32()
        N.check(self._lib.olhip_engine_step(self.handle, N.ctypes.byref(pending)))
        return int(pending.value)

    def run(self):
        """Step until every submitted job is done."""
        N.check(self._lib.olhip_engine_run(self.handle))

    def poll(self, job_id: int) -> int:
        state = N.ctypes.c_int()
        N.check(self._lib.olhip_engine_poll(self.handle, int(job_id), N.ctypes.byref(state)))
        return int(state.value)

    def wait(self, job_id: int, timeout: float | None = None) -> int:
        state = N.ctypes.c_int()
        N.check(self._lib.olhip_engine_wait(self.handle, int(job_id), -1.0 if timeout is None else float(timeout),
                                            N.ctypes.byref(state)))
        return int(state.value)

    def result(self, job_id: int, copy_image=True):
        """(float32 (H, W, 4) image or None, statistics dict); releases the job."""
        job = self._jobs.pop(int(job_id), None)
        if job is None:
            raise ValueError("unknown job id")
        out = np.empty((job.height, job.width, 4), np.float32) if copy_image else None
        stats = N.JobStats()
        N.check(self._lib.olhip_engine_result(self.handle, int(job_id), N.ptr(out) if out is not None else None,
                                              N.ctypes.byref(stats)))
        return out, stats.as_dict()

    def products(self, job_id: int):
        """Named products of a done job (call before :meth:`result`): dict of
        depth, normal, primitive (primary products) and variance."""
        job = self._jobs[int(job_id)]
        h, w = job.height, job.width
        out = {}
        if job.products & N.PRODUCT_PRIMARY:
            out["depth"] = np.empty((h, w), np.float32)
            out["normal"] = np.empty((h, w, 3), np.float32)
            out["primitive"] = np.empty((h, w), np.uint32)
        if job.products & N.PRODUCT_VARIANCE:
            out["variance"] = np.empty((h, w), np.float32)
        N.check(self._lib.olhip_engine_products(
            self.handle, int(job_id), *(N.ptr(out[k], N.ctypes.c_uint32 if k == "primitive" else N.ctypes.c_float)
                                        if k in out else None for k in ("depth", "normal", "primitive", "variance"))))
        return out

    def start(self):
        N.check(self._lib.olhip_engine_start(self.handle))
        self._worker = True

    def stop(self):
        if self._worker:
            N.check(self._lib.olhip_engine_stop(self.handle))
            self._worker = False

    def stats(self) -> dict:
        s = N.EngineStats()
        N.check(self._lib.olhip_engine_get_stats(self.handle, N.ctypes.byref(s)))
        return s.as_dict()

    def close(self):
        if self.handle:
            self.stop()
            N.check(self._lib.olhip_engine_destroy(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def plain_render(job: Job, copy_image=True):
    """One job through the single-frame entry point (``olhip_render_gi``)."""
    lib = N.library()
    ctx = job.scene.context
    position, forward, right, up = camera_basis(job.camera)
    cam = N.Camera()
    cam.position[:] = position.tolist()
    cam.forward[:] = forward.tolist()
    cam.right[:] = right.tolist()
    cam.up[:] = up.tolist()
    cam.projection = camera_projection(job.camera)
    settings = job.settings()
    out = np.empty((job.height, job.width, 4), np.float32) if copy_image else None
    stats = N.Stats()
    N.check(lib.olhip_render_gi(ctx.handle, job.scene.handle, N.ctypes.byref(cam), int(job.width),
                                int(job.height), int(job.samples), int(job.frame_index),
                                N.ctypes.byref(settings), N.ptr(out) if out is not None else None,
                                N.ctypes.byref(stats)))
    if stats.queue_overflow:
        raise RuntimeError(f"ray queue overflow: {stats.queue_overflow} continuations or shadow groups dropped")
    return out, stats.as_dict()


def run_batch(engine: BatchEngine, jobs, copy_images=False):
    """Submit every job, run to completion; returns (wall seconds, results)."""
    started = time.perf_counter()
    ids = [engine.submit(j) for j in jobs]
    engine.run()
    wall = time.perf_counter() - started
    return wall, [engine.result(i, copy_image=copy_images) for i in ids]


def autotune(context: DeviceContext, jobs, *, slots=(1 << 22, 1 << 23, 1 << 24, 1 << 25, 1 << 26, 1 << 27),
             shapes=((16, 128), (8, 128), (16, 256), (32, 128)), repeats=1, measure=None):
    """Choose path-slot capacity and (LDS stack entries, block size) for a job mix.

    Coordinate search: the launch shape at the largest capacity first, then
    the capacity at the best shape, each candidate timed on the same
    calibration jobs (rays per second over the whole batch).  ``measure``
    (settings -> seconds) replaces the timing, for tests.  Returns
    (best EngineSettings, {settings: seconds}).
    """
    timings = {}

    def timed(settings):
        if settings in timings:
            return timings[settings]
        if measure is not None:
            seconds = float(measure(settings))
        else:
            best = math.inf
            with BatchEngine(context, settings) as engine:
                run_batch(engine, jobs[: max(1, min(len(jobs), 8))])  # warm-up: graphs, buffers
                for _ in range(repeats):
                    best = min(best, run_batch(engine, jobs)[0])
            seconds = best
        timings[settings] = seconds
        return seconds

    largest = max(slots)
    shape = min(shapes, key=lambda s: timed(EngineSettings(path_slots=largest, lds_stack=s[0], trace_block=s[1])))
    size = min(slots, key=lambda n: timed(EngineSettings(path_slots=n, lds_stack=shape[0], trace_block=shape[1])))
    return EngineSettings(path_slots=size, lds_stack=shape[0], trace_block=shape[1]), timings


def poisson_arrivals(rate, count, seed=0):
    """Arrival times (seconds from 0) of a Poisson process of the given rate."""
    if rate <= 0:
        raise ValueError("rate must be positive")
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=int(count)))


def latency_percentiles(latencies_ms):
    """p50 / p99 / mean of a list of latencies (nearest-rank percentiles)."""
    lat = np.sort(np.asarray(latencies_ms, np.float64))
    if lat.size == 0:
        raise ValueError("no latencies")

    def rank(p):
        return float(lat[min(lat.size - 1, max(0, int(math.ceil(p / 100.0 * lat.size)) - 1))])

    return {"p50_ms": rank(50), "p99_ms": rank(99), "mean_ms": float(lat.mean()), "count": int(lat.size)}


class RenderServer:
    """Continuous-batching server over resident scenes.

    Scenes are uploaded once (BVH + materials stay in HBM).  ``submit`` is
    thread-safe and returns immediately; the engine's worker thread admits
    queued jobs as path and shadow-entry slots free up (admission by memory)
    and advances all resident jobs together.
    """

    def __init__(self, context: DeviceContext, settings: EngineSettings | None = None):
        self.context = context
        self.engine = BatchEngine(context, settings)
        self.scenes = {}
        self._lock = threading.Lock()
        self._ids = itertools.count()
        self.engine.start()

    def add_scene(self, key, scene, leaf_size=2):
        self.scenes[key] = DeviceScene(self.context, scene, leaf_size)
        return self.scenes[key]

    def submit(self, key, camera, width, height, samples=1, max_bounces=5, frame_index=0, config=None):
        job = Job(self.scenes[key], camera, width, height, samples, frame_index, config, max_bounces)
        with self._lock:
            return self.engine.submit(job)

    def wait(self, job_id, timeout=None, copy_image=True):
        state = self.engine.wait(job_id, timeout)
        if state != N.JOB_DONE:
            if state == N.JOB_FAILED:
                self.engine.result(job_id)  # raises with the reason
            raise TimeoutError(f"job {job_id} not done (state {state})")
        with self._lock:
            return self.engine.result(job_id, copy_image=copy_image)

    def close(self):
        self.engine.close()
        for s in self.scenes.values():
            s.close()
        self.scenes.clear()
