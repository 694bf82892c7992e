"""Camera models (``ordinarylight/cameras/*.py``) and the camera basis.

``camera_basis`` restates ``_camera_vectors`` / ``_camera_projection``
(``ordinarylight/targets/vulkan/core.py:95-121``): float32 forward/right/up
with right and up pre-scaled by the projection's half extent.  The GI kernels
consume exactly this basis, so the same float32 arithmetic is kept.
"""

from __future__ import annotations

from dataclasses import dataclass
import math

import numpy as np


def _validate_look_at(position, target, up):
    position = np.asarray(position, dtype=np.float64)
    target = np.asarray(target, dtype=np.float64)
    up = np.asarray(up, dtype=np.float64)
    for name, value in (("position", position), ("target", target), ("up", up)):
        if value.shape != (3,) or not np.all(np.isfinite(value)):
            raise ValueError(f"camera {name} must contain three finite values")
    forward = target - position
    if np.linalg.norm(forward) < 1e-8:
        raise ValueError("camera position and target must differ")
    if np.linalg.norm(np.cross(forward, up)) < 1e-8:
        raise ValueError("camera up vector cannot be parallel to the view direction")


@dataclass(frozen=True)
class PerspectiveCamera:
    """Pinhole camera described by a look-at transform and vertical FOV."""

    position: tuple
    target: tuple
    up: tuple = (0.0, 1.0, 0.0)
    vertical_fov_degrees: float = 45.0

    def __post_init__(self):
        _validate_look_at(self.position, self.target, self.up)
        if not 1.0 <= self.vertical_fov_degrees < 179.0:
            raise ValueError("vertical_fov_degrees must be in [1, 179)")


@dataclass(frozen=True)
class OrthographicCamera:
    """Parallel projection with a world-space vertical view size."""

    position: tuple
    target: tuple
    up: tuple = (0.0, 1.0, 0.0)
    vertical_size: float = 2.0

    def __post_init__(self):
        _validate_look_at(self.position, self.target, self.up)
        if not math.isfinite(self.vertical_size) or self.vertical_size <= 0.0:
            raise ValueError("vertical_size must be finite and positive")

    @property
    def vertical_fov_degrees(self):
        return 90.0


@dataclass(frozen=True)
class PanoramicCamera:
    """Equirectangular camera centred on its look-at orientation."""

    position: tuple
    target: tuple
    up: tuple = (0.0, 1.0, 0.0)
    horizontal_fov_degrees: float = 360.0
    vertical_fov_degrees: float = 180.0

    def __post_init__(self):
        _validate_look_at(self.position, self.target, self.up)
        if not 1.0 <= self.horizontal_fov_degrees <= 360.0:
            raise ValueError("horizontal_fov_degrees must be in [1, 360]")
        if not 1.0 <= self.vertical_fov_degrees <= 180.0:
            raise ValueError("vertical_fov_degrees must be in [1, 180]")


CAMERA_TYPES = (PerspectiveCamera, OrthographicCamera, PanoramicCamera)
Camera = PerspectiveCamera | OrthographicCamera | PanoramicCamera

PROJECTION_PERSPECTIVE, PROJECTION_ORTHOGRAPHIC, PROJECTION_PANORAMIC = 0, 1, 2


def camera_projection(camera):
    """Projection code: 0 perspective, 1 orthographic, 2 panoramic."""
    name = type(camera).__name__
    if name == "OrthographicCamera":
        return PROJECTION_ORTHOGRAPHIC
    if name == "PanoramicCamera":
        return PROJECTION_PANORAMIC
    return PROJECTION_PERSPECTIVE


def _cross3(a, b):
    """float32 cross product of two 3-vectors, rounded as ``np.cross`` rounds
    (each product, then the difference), without its per-call overhead."""
    a0, a1, a2 = a
    b0, b1, b2 = b
    return np.array((a1 * b2 - a2 * b1, a2 * b0 - a0 * b2, a0 * b1 - a1 * b0), dtype=np.float32)


def camera_basis(camera):
    """float32 ``(position, forward, right, up)`` with scaled right/up."""
    position = np.asarray(camera.position, dtype=np.float32)
    forward = np.asarray(camera.target, dtype=np.float32) - position
    forward /= np.linalg.norm(forward)
    right = _cross3(forward, np.asarray(camera.up, dtype=np.float32))
    right /= np.linalg.norm(right)
    up = _cross3(right, forward)
    projection = camera_projection(camera)
    if projection == PROJECTION_ORTHOGRAPHIC:
        scale = float(camera.vertical_size) * 0.5
        right *= scale
        up *= scale
    elif projection == PROJECTION_PANORAMIC:
        right *= math.radians(camera.horizontal_fov_degrees) * 0.5
        up *= math.radians(camera.vertical_fov_degrees) * 0.5
    else:
        scale = math.tan(math.radians(camera.vertical_fov_degrees) * 0.5)
        right *= scale
        up *= scale
    return position, forward, right, up


__all__ = [
    "CAMERA_TYPES", "Camera", "OrthographicCamera", "PanoramicCamera",
    "PerspectiveCamera", "camera_basis", "camera_projection",
]
