"""High-level ``Renderer`` with the reference's frame-index and job semantics.

Mirrors ``ordinarylight/renderer/_core.py``: ``render`` / ``render_async``
return float32 ``(H, W, 4)`` linear HDR; frame indices are reserved at
submission (``:639-646``) so cancelled jobs leave gaps; all calls run in
order on one worker thread (``:263-265``) under an implementation lock
(``:712``); ``reset_sequence``, ``replace_scene`` and ``reconfigure`` are
ordered with renders; ``out=`` reuses a caller array; named products are
validated against the reference's dtype/shape table (``:768-784``).

The default implementation is :class:`~ordinarylight_amd.renderers.hip.HipPathTracer`.
"""

from __future__ import annotations

import asyncio
from collections.abc import Iterator, Mapping
from concurrent.futures import ThreadPoolExecutor, wait
from dataclasses import dataclass
from threading import Lock
from types import MappingProxyType
from typing import Any

import numpy as np

from .cameras import CAMERA_TYPES
from .capabilities import capabilities_from_renderer

_EXPECTED = {
    "color": (3, np.dtype(np.float32), 4),
    "variance": (2, np.dtype(np.float32), None),
    "depth": (2, np.dtype(np.float32), None),
    "normal": (3, np.dtype(np.float32), 3),
    "instance_id": (2, np.dtype(np.uint32), None),
    "object_id": (2, np.dtype(np.uint32), None),
    "material_id": (2, np.dtype(np.uint32), None),
    "motion": (3, np.dtype(np.float32), 2),
}


def _expected_shape(name, width, height):
    ndim, dtype, last = _EXPECTED[name]
    return ((height, width) if ndim == 2 else (height, width, last)), dtype


class RenderStatistics(Mapping):
    """Immutable measurements for one completed render."""

    def __init__(self, *, frame_index, size, samples, timings=None):
        timings = dict(timings or {})
        self._timings = MappingProxyType(timings)
        self._values = MappingProxyType({
            "frame_index": int(frame_index), "width": int(size[0]), "height": int(size[1]),
            "samples": samples,
            "total_ms": self._duration(timings, "total_ms", "frame_ms", "cpu_frame_ms"),
            "gpu_ms": self._duration(timings, "gpu_ms", "gpu_frame_ms", "dispatch_gpu_ms"),
        })

    @staticmethod
    def _duration(timings, *names):
        for name in names:
            value = timings.get(name)
            if value is not None:
                return float(value)
        return None

    @property
    def timings(self):
        return self._timings

    def as_dict(self, *, include_timings=True):
        record = dict(self._values)
        if include_timings:
            for name, value in self._timings.items():
                record.setdefault(name, value)
        return record

    def __getitem__(self, name):
        return self._values[name]

    def __iter__(self) -> Iterator[str]:
        return iter(self._values)

    def __len__(self):
        return len(self._values)

    def __getattr__(self, name):
        try:
            return self._values[name]
        except KeyError as error:
            raise AttributeError(name) from error


class RenderFrame(Mapping):
    """Immutable named collection of products plus frame metadata."""

    def __init__(self, products, *, metadata=None):
        if not products:
            raise ValueError("a render frame requires at least one product")
        for name, value in products.items():
            if not isinstance(name, str) or not name:
                raise TypeError("render product names must be non-empty strings")
            if not isinstance(value, np.ndarray):
                raise TypeError(f"render product {name!r} must be a NumPy array")
        self._products = MappingProxyType(dict(products))
        self._metadata = MappingProxyType(dict(metadata or {}))

    @property
    def metadata(self):
        return self._metadata

    def __getitem__(self, name):
        return self._products[name]

    def __iter__(self):
        return iter(self._products)

    def __len__(self):
        return len(self._products)

    def __getattr__(self, name):
        try:
            return self._products[name]
        except KeyError as error:
            raise AttributeError(name) from error


@dataclass(frozen=True)
class _CompletedRender:
    value: Any
    statistics: RenderStatistics


class RenderJob:
    """A nonblocking submission; jobs of one renderer execute in order."""

    def __init__(self, future, *, frame_index, scene, camera):
        self._future = future
        self._frame_index = int(frame_index)
        self._scene = scene
        self._camera = camera

    @property
    def frame_index(self):
        return self._frame_index

    def done(self):
        return self._future.done()

    def running(self):
        return self._future.running()

    def cancelled(self):
        return self._future.cancelled()

    @property
    def statistics(self):
        if not self.done() or self.cancelled():
            return None
        try:
            return self._future.result().statistics
        except Exception:
            return None

    def cancel(self):
        return self._future.cancel()

    def wait(self, timeout=None):
        done, _ = wait((self._future,), timeout=timeout)
        return bool(done)

    def result(self, timeout=None):
        return self._future.result(timeout=timeout).value

    def exception(self, timeout=None):
        return self._future.exception(timeout=timeout)

    def add_done_callback(self, callback):
        if not callable(callback):
            raise TypeError("callback must be callable")
        self._future.add_done_callback(lambda _f: callback(self))
        return self

    async def _async_result(self):
        completed = await asyncio.wrap_future(self._future)
        return completed.value

    def __await__(self):
        return self._async_result().__await__()


def _size(value):
    try:
        width, height = value
    except (TypeError, ValueError) as error:
        raise TypeError("size must be a (width, height) pair") from error
    width, height = int(width), int(height)
    if width < 1 or height < 1:
        raise ValueError("render dimensions must be positive")
    return width, height


def _default_implementation(config, options):
    from .renderers.hip import HipPathTracer, HipRendererConfig
    if config is not None and not isinstance(config, HipRendererConfig):
        raise TypeError("config must be a HipRendererConfig")
    return HipPathTracer(config) if config is not None else HipPathTracer(**options)


class Renderer:
    """Render scenes to HDR NumPy arrays; the HIP GI path tracer by default."""

    def __init__(self, *, config=None, implementation=None, **config_options):
        if config is not None and config_options:
            raise TypeError("pass config or renderer options, not both")
        if implementation is not None and (config is not None or config_options):
            raise TypeError("a supplied implementation owns its configuration")
        self._implementation = (implementation if implementation is not None
                                else _default_implementation(config, config_options))
        if not callable(getattr(self._implementation, "render_frame", None)):
            raise TypeError("implementation must implement RendererProtocol.render_frame()")
        if not callable(getattr(self._implementation, "close", None)):
            raise TypeError("implementation must provide close()")
        self._frame_index = 0
        self._last_statistics = None
        self._capabilities = capabilities_from_renderer(self._implementation)
        self._state_lock = Lock()
        self._submission_lock = Lock()
        self._implementation_lock = Lock()
        self._accepting_jobs = True
        self._executor = ThreadPoolExecutor(max_workers=1, thread_name_prefix="ordinarylight-render")

    @property
    def config(self):
        return getattr(self._implementation, "config", None)

    @property
    def device(self):
        return getattr(self._implementation, "device", None)

    @property
    def last_timings(self):
        return dict(getattr(self._implementation, "last_timings", {}))

    @property
    def last_statistics(self):
        return self._last_statistics

    @property
    def frame_index(self):
        with self._state_lock:
            return self._frame_index

    @property
    def available_outputs(self):
        if self._implementation is None:
            return ()
        return self._capabilities.outputs

    @property
    def capabilities(self):
        return self._capabilities

    def reset_sequence(self, frame_index=0):
        frame_index = int(frame_index)
        if frame_index < 0:
            raise ValueError("frame_index cannot be negative")
        with self._submission_lock:
            with self._state_lock:
                if not self._accepting_jobs:
                    raise RuntimeError("renderer is closed")
                self._frame_index = frame_index
            self._executor.submit(self._reset_history).result()

    def _reset_history(self):
        with self._implementation_lock:
            reset = getattr(self._implementation, "reset_output_history", None)
            if callable(reset):
                reset()

    def replace_scene(self, scene, *, reset_sequence=True):
        with self._submission_lock:
            with self._state_lock:
                if not self._accepting_jobs or self._implementation is None:
                    raise RuntimeError("renderer is closed")
            self._executor.submit(self._replace_scene_serialized, scene).result()
            if reset_sequence:
                with self._state_lock:
                    self._frame_index = 0
        return scene

    def _replace_scene_serialized(self, scene):
        with self._implementation_lock:
            replace = getattr(self._implementation, "replace_scene", None)
            if not callable(replace):
                raise RuntimeError("the active renderer implementation does not support "
                                   "resident scene replacement")
            replace(scene)
            self._last_statistics = None

    def reconfigure(self, **changes):
        if not changes:
            return self.config
        with self._submission_lock:
            with self._state_lock:
                if not self._accepting_jobs or self._implementation is None:
                    raise RuntimeError("renderer is closed")
            return self._executor.submit(self._reconfigure_serialized, dict(changes)).result()

    def _reconfigure_serialized(self, changes):
        with self._implementation_lock:
            reconfigure = getattr(self._implementation, "reconfigure", None)
            if not callable(reconfigure):
                raise RuntimeError("the active renderer implementation requires renderer recreation for settings")
            config = reconfigure(**changes)
            self._last_statistics = None
            return config

    def render(self, scene, camera, size, *, samples=None, frame_index=None, out=None, outputs=None):
        """Render one frame (blocking); advances the sequence when frame_index is None."""
        return self.render_async(scene, camera, size, samples=samples, frame_index=frame_index,
                                 out=out, outputs=outputs).result()

    def render_async(self, scene, camera, size, *, samples=None, frame_index=None, out=None, outputs=None):
        if not isinstance(camera, CAMERA_TYPES) and type(camera).__name__ not in (
                "PerspectiveCamera", "OrthographicCamera", "PanoramicCamera"):
            raise TypeError("camera must be a supported ordinarylight camera")
        width, height = _size(size)
        with self._submission_lock:
            with self._state_lock:
                if not self._accepting_jobs or self._implementation is None:
                    raise RuntimeError("renderer is closed")
                current = self._frame_index if frame_index is None else int(frame_index)
                if current < 0:
                    raise ValueError("frame_index cannot be negative")
                self._frame_index = current + 1
            future = self._executor.submit(self._render_serialized, scene, camera, width, height,
                                           samples, current, out, outputs)
        return RenderJob(future, frame_index=current, scene=scene, camera=camera)

    def _render_serialized(self, scene, camera, width, height, samples, current, out, outputs):
        with self._implementation_lock:
            value = self._render_now(scene, camera, width, height, samples=samples,
                                     current_frame=current, out=out, outputs=outputs)
            return _CompletedRender(value, self._last_statistics)

    def _render_now(self, scene, camera, width, height, *, samples, current_frame, out, outputs):
        named = outputs is not None
        if outputs is None:
            requested = ("color",)
        elif isinstance(outputs, str):
            requested = (outputs,)
        else:
            requested = tuple(outputs)
        if not requested or any(not isinstance(n, str) or not n for n in requested):
            raise ValueError("outputs must contain one or more product names")
        if len(set(requested)) != len(requested):
            raise ValueError("outputs cannot contain duplicate names")
        unavailable = tuple(n for n in requested if n not in self.available_outputs)
        if unavailable:
            raise ValueError(f"unsupported render outputs {unavailable}; available outputs "
                             f"are {self.available_outputs}")
        if named and out is not None and not isinstance(out, Mapping):
            raise TypeError("named outputs require out to be a mapping of arrays")
        if not named and out is not None:
            if not isinstance(out, np.ndarray):
                raise TypeError("out must be a NumPy array")
            if out.shape != (height, width, 4):
                raise ValueError(f"out must have shape {(height, width, 4)}, got {out.shape}")
            if out.dtype != np.float32:
                raise TypeError("out must use float32 components")
        if named and callable(getattr(self._implementation, "render_products", None)):
            products = self._implementation.render_products(
                scene, camera, width, height, outputs=requested, samples=samples, frame_index=current_frame)
        else:
            products = {"color": self._implementation.render_frame(
                scene, camera, width, height, samples=samples, frame_index=current_frame)}
        products = {name: np.asarray(products[name]) for name in requested}
        for name, result in products.items():
            shape, dtype = _expected_shape(name, width, height)
            if result.shape != shape or result.dtype != dtype:
                raise RuntimeError(f"renderer returned invalid {name!r}; expected {dtype} {shape}, "
                                   f"got {result.dtype} {result.shape}")
        statistics = RenderStatistics(frame_index=current_frame, size=(width, height), samples=samples,
                                      timings=self.last_timings)
        self._last_statistics = statistics
        if named:
            if out is not None:
                unknown = set(out) - set(requested)
                if unknown:
                    raise ValueError(f"out contains unrequested products: {unknown}")
                for name, destination in out.items():
                    if not isinstance(destination, np.ndarray):
                        raise TypeError(f"out[{name!r}] must be a NumPy array")
                    if destination.shape != products[name].shape or destination.dtype != products[name].dtype:
                        raise ValueError(f"out[{name!r}] must match {products[name].dtype} {products[name].shape}")
                    np.copyto(destination, products[name])
                    products[name] = destination
            return RenderFrame(products, metadata={
                "frame_index": current_frame, "size": (width, height), "samples": samples,
                "timings": self.last_timings, "statistics": statistics})
        result = products["color"]
        if out is not None:
            np.copyto(out, result)
            return out
        return result

    def close(self):
        with self._submission_lock:
            with self._state_lock:
                if not self._accepting_jobs:
                    return
                self._accepting_jobs = False
        self._executor.shutdown(wait=True, cancel_futures=True)
        with self._implementation_lock:
            if self._implementation is not None:
                self._implementation.close()
                self._implementation = None

    def __enter__(self):
        if self._implementation is None:
            raise RuntimeError("renderer is closed")
        return self

    def __exit__(self, *_args):
        self.close()


__all__ = ["RenderFrame", "RenderJob", "RenderStatistics", "Renderer"]
