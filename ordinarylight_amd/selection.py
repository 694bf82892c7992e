"""Picking types: the options, viewport mapping and result of
``ordinarylight/selection/_core.py`` (``PickOptions`` :15-37,
``ViewportMapping`` :40-86, ``PickResult`` :89-99), same fields and checks."""

from __future__ import annotations

from dataclasses import dataclass
import math

import numpy as np


def _size(value, name):
    if len(value) != 2:
        raise ValueError(f"{name} must contain width and height")
    out = tuple(int(v) for v in value)
    if min(out) < 1:
        raise ValueError(f"{name} dimensions must be positive")
    return out


@dataclass(frozen=True)
class PickOptions:
    """Which intersections may be selected: ``transmissive`` "surface" or
    "through", the transmission threshold, and ``volumes`` "include",
    "exclude" or "only"."""

    transmissive: str = "surface"
    transmission_threshold: float = 0.001
    volumes: str = "include"

    def __post_init__(self):
        if self.transmissive not in ("surface", "through"):
            raise ValueError("transmissive must be 'surface' or 'through'")
        threshold = float(self.transmission_threshold)
        if not math.isfinite(threshold) or not 0.0 <= threshold <= 1.0:
            raise ValueError("transmission_threshold must be in [0, 1]")
        if self.volumes not in ("include", "exclude", "only"):
            raise ValueError("volumes must be 'include', 'exclude', or 'only'")
        object.__setattr__(self, "transmission_threshold", threshold)


@dataclass(frozen=True)
class ViewportMapping:
    """UI pixel -> render pixel through a displayed content rectangle
    ``(x, y, width, height)``, a framebuffer scale and a render size."""

    viewport_size: tuple
    framebuffer_size: tuple | None = None
    render_size: tuple | None = None
    content_rect: tuple | None = None

    def __post_init__(self):
        viewport = _size(self.viewport_size, "viewport_size")
        framebuffer = _size(self.framebuffer_size or viewport, "framebuffer_size")
        render = _size(self.render_size or framebuffer, "render_size")
        rect = tuple(float(v) for v in (self.content_rect or (0.0, 0.0, float(viewport[0]), float(viewport[1]))))
        if len(rect) != 4 or not np.isfinite(rect).all():
            raise ValueError("content_rect must contain four finite values")
        if rect[2] <= 0.0 or rect[3] <= 0.0:
            raise ValueError("content_rect dimensions must be positive")
        object.__setattr__(self, "viewport_size", viewport)
        object.__setattr__(self, "framebuffer_size", framebuffer)
        object.__setattr__(self, "render_size", render)
        object.__setattr__(self, "content_rect", rect)

    def map_pixel(self, pixel, *, target="render"):
        """The mapped pixel, or None outside the displayed content."""
        if target not in ("framebuffer", "render"):
            raise ValueError("target must be 'framebuffer' or 'render'")
        x, y = (float(v) for v in pixel)
        if not np.isfinite((x, y)).all():
            raise ValueError("pixel coordinates must be finite")
        left, top, width, height = self.content_rect
        if x < left or y < top or x >= left + width or y >= top + height:
            return None
        size = self.framebuffer_size if target == "framebuffer" else self.render_size
        return (x - left) / width * size[0], (y - top) / height * size[1]


@dataclass(frozen=True)
class PickResult:
    """The closest selectable object hit by a pixel's primary ray."""

    object: object
    object_id: int
    distance: float
    position: tuple
    primitive_index: int
    triangle_index: int
    barycentric: tuple
