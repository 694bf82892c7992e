"""Picking records of the reference's selection API.

The field names and meanings are the interface ``HipPathTracer.pick``
answers to (``ordinarylight/selection/_core.py``: ``PickOptions`` :15-37,
``ViewportMapping`` :40-86, ``PickResult`` :89-99).  A viewport pixel maps
through the displayed content rectangle onto the framebuffer or render
grid; everything else here is validation of those records.
"""

from __future__ import annotations

from dataclasses import dataclass
import math

_TRANSMISSIVE = ("surface", "through")
_VOLUMES = ("include", "exclude", "only")


def _two_positive_ints(value, what):
    pair = tuple(value)
    if len(pair) != 2:
        raise ValueError(f"{what}: expected (width, height), got {len(pair)} values")
    pair = tuple(int(v) for v in pair)
    if pair[0] < 1 or pair[1] < 1:
        raise ValueError(f"{what}: width and height must be at least 1, got {pair}")
    return pair


def _finite(values, what):
    out = tuple(float(v) for v in values)
    if not all(math.isfinite(v) for v in out):
        raise ValueError(f"{what}: values must be finite, got {out}")
    return out


@dataclass(frozen=True)
class PickOptions:
    """Which surfaces a pick may return: glass like any surface
    (``transmissive="surface"``) or skipped above ``transmission_threshold``
    (``"through"``); volume proxies included, excluded or alone."""

    transmissive: str = "surface"
    transmission_threshold: float = 0.001
    volumes: str = "include"

    def __post_init__(self):
        if self.transmissive not in _TRANSMISSIVE:
            raise ValueError(f"transmissive: one of {_TRANSMISSIVE}, got {self.transmissive!r}")
        if self.volumes not in _VOLUMES:
            raise ValueError(f"volumes: one of {_VOLUMES}, got {self.volumes!r}")
        (threshold,) = _finite((self.transmission_threshold,), "transmission_threshold")
        if threshold < 0.0 or threshold > 1.0:
            raise ValueError(f"transmission_threshold: expected a value in [0, 1], got {threshold}")
        object.__setattr__(self, "transmission_threshold", threshold)


@dataclass(frozen=True)
class ViewportMapping:
    """A UI viewport showing the render inside ``content_rect`` (x, y, width,
    height in viewport pixels, e.g. letterboxed); the framebuffer may be
    scaled for high DPI and the render grid may differ again.  Omitted sizes
    default to the previous one in that chain."""

    viewport_size: tuple
    framebuffer_size: tuple | None = None
    render_size: tuple | None = None
    content_rect: tuple | None = None

    def __post_init__(self):
        viewport = _two_positive_ints(self.viewport_size, "viewport_size")
        framebuffer = _two_positive_ints(self.framebuffer_size or viewport, "framebuffer_size")
        render = _two_positive_ints(self.render_size or framebuffer, "render_size")
        rect = self.content_rect if self.content_rect is not None else (0, 0, viewport[0], viewport[1])
        rect = _finite(rect, "content_rect")
        if len(rect) != 4:
            raise ValueError(f"content_rect: expected (x, y, width, height), got {len(rect)} values")
        if min(rect[2], rect[3]) <= 0.0:
            raise ValueError(f"content_rect: width and height must be positive, got {rect[2:]}")
        for name, value in (("viewport_size", viewport), ("framebuffer_size", framebuffer),
                            ("render_size", render), ("content_rect", rect)):
            object.__setattr__(self, name, value)

    def map_pixel(self, pixel, *, target="render"):
        """``pixel`` (viewport coordinates) on the ``target`` grid
        ("render" or "framebuffer"), or None outside the content rectangle."""
        grids = {"render": self.render_size, "framebuffer": self.framebuffer_size}
        if target not in grids:
            raise ValueError(f"target: one of {tuple(grids)}, got {target!r}")
        px, py = _finite(pixel, "pixel")
        x0, y0, w, h = self.content_rect
        inside = x0 <= px < x0 + w and y0 <= py < y0 + h
        if not inside:
            return None
        gw, gh = grids[target]
        return (px - x0) / w * gw, (py - y0) / h * gh


@dataclass(frozen=True)
class PickResult:
    """The closest selectable surface under a pixel: the scene object, its
    id, the distance and position along the pixel's primary ray, the global
    and per-object triangle index and the hit's barycentric weights."""

    object: object
    object_id: int
    distance: float
    position: tuple
    primitive_index: int
    triangle_index: int
    barycentric: tuple
