"""Renderer contracts (``ordinarylight/renderers/base.py:16-143``).

Structural protocols identical in meaning to the reference's, with ``"hip"``
added to the graphics-API vocabulary for the MI355X implementation.
"""

from __future__ import annotations

from collections.abc import Mapping
from dataclasses import dataclass
from typing import Any, ClassVar, Literal, Protocol, runtime_checkable

import numpy as np

RendererFamily = Literal["global_illumination", "raster", "hybrid", "reference"]
GraphicsApi = Literal["vulkan", "webgpu", "cpu", "composite", "hip"]


@dataclass(frozen=True)
class RendererImplementationInfo:
    """Stable identity for a concrete renderer implementation."""

    name: str
    family: RendererFamily
    graphics_api: GraphicsApi


class RendererImplementation:
    """Nominal base carrying an implementation identity."""

    implementation: ClassVar[RendererImplementationInfo]

    @classmethod
    def implementation_info(cls) -> RendererImplementationInfo:
        return cls.implementation


@runtime_checkable
class RendererProtocol(Protocol):
    """``render_frame`` returns float32 (H, W, 4) linear HDR; ``close`` is idempotent."""

    available_outputs: tuple[str, ...]
    config: Any
    device: Any
    last_timings: Mapping[str, Any]

    def render_frame(self, scene, camera, width: int, height: int, *,
                     samples: int | None = None, frame_index: int = 0) -> np.ndarray:
        ...

    def close(self) -> None:
        ...


@runtime_checkable
class ProductRendererProtocol(RendererProtocol, Protocol):
    def render_products(self, scene, camera, width: int, height: int, *,
                        outputs: tuple[str, ...], samples: int | None = None,
                        frame_index: int = 0) -> Mapping[str, np.ndarray]:
        ...


@runtime_checkable
class ResidentSceneRendererProtocol(RendererProtocol, Protocol):
    def replace_scene(self, scene) -> None:
        ...


__all__ = [
    "GraphicsApi", "ProductRendererProtocol", "RendererFamily",
    "RendererImplementation", "RendererImplementationInfo", "RendererProtocol",
    "ResidentSceneRendererProtocol",
]
