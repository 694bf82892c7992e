"""Analytic lights and the environment light.

Field names and validation follow ``ordinarylight/lights/*.py``
(``environment_light.py:107-133`` for the environment).
"""

from __future__ import annotations

from dataclasses import dataclass, field
import math

import numpy as np


def _color_intensity(color, intensity):
    color = np.asarray(color, dtype=np.float64)
    if color.shape != (3,) or not np.all(np.isfinite(color)) or np.any(color < 0.0):
        raise ValueError("light color must contain three finite non-negative values")
    if not math.isfinite(float(intensity)) or float(intensity) < 0.0:
        raise ValueError("light intensity must be finite and non-negative")


def _vec3(value, name):
    array = np.asarray(value, dtype=np.float64)
    if array.shape != (3,) or not np.all(np.isfinite(array)):
        raise ValueError(f"{name} must contain three finite values")
    return array


def _range(value):
    if value is not None and (not math.isfinite(float(value)) or float(value) <= 0.0):
        raise ValueError("light range must be positive and finite or None")


@dataclass(eq=False)
class PointLight:
    position: tuple
    color: tuple = (1.0, 1.0, 1.0)
    intensity: float = 1.0
    range: float | None = None
    id: int | None = field(default=None, init=False)

    def __post_init__(self):
        _vec3(self.position, "position")
        _color_intensity(self.color, self.intensity)
        _range(self.range)


@dataclass(eq=False)
class DirectionalLight:
    direction: tuple
    color: tuple = (1.0, 1.0, 1.0)
    intensity: float = 1.0
    id: int | None = field(default=None, init=False)

    def __post_init__(self):
        if np.linalg.norm(_vec3(self.direction, "direction")) < 1e-8:
            raise ValueError("direction cannot be zero")
        _color_intensity(self.color, self.intensity)


@dataclass(eq=False)
class SpotLight:
    position: tuple
    direction: tuple
    color: tuple = (1.0, 1.0, 1.0)
    intensity: float = 1.0
    range: float | None = None
    inner_cone_angle: float = 0.3
    outer_cone_angle: float = 0.5
    id: int | None = field(default=None, init=False)

    def __post_init__(self):
        _vec3(self.position, "position")
        if np.linalg.norm(_vec3(self.direction, "direction")) < 1e-8:
            raise ValueError("direction cannot be zero")
        _color_intensity(self.color, self.intensity)
        _range(self.range)
        if not 0.0 <= self.inner_cone_angle <= self.outer_cone_angle <= math.pi / 2:
            raise ValueError("cone angles must satisfy 0 <= inner <= outer <= pi/2")


@dataclass(frozen=True)
class EnvironmentLight:
    """Uniform (optionally image-based) distant radiance."""

    image: np.ndarray | None = field(default=None, compare=False, repr=False)
    color: tuple = (1.0, 1.0, 1.0)
    intensity: float = 1.0
    rotation: float = 0.0

    def __post_init__(self):
        _color_intensity(self.color, self.intensity)
        if not math.isfinite(self.rotation):
            raise ValueError("environment rotation must be finite")
        if self.image is None:
            return
        image = np.array(self.image, dtype=np.float32, copy=True, order="C")
        if image.ndim != 3 or image.shape[2] not in (3, 4):
            raise ValueError("environment image must have shape (height, width, 3 or 4)")
        image = np.ascontiguousarray(image[..., :3], dtype=np.float32)
        if not np.all(np.isfinite(image)) or np.any(image < 0.0):
            raise ValueError("environment image radiance must be finite and non-negative")
        image.flags.writeable = False
        object.__setattr__(self, "image", image)


LIGHT_TYPES = (PointLight, DirectionalLight, SpotLight)

__all__ = ["DirectionalLight", "EnvironmentLight", "LIGHT_TYPES", "PointLight", "SpotLight"]
