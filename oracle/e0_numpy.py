"""O-E0: NumPy restatement of OrdinaryLight's CPU reference path tracer.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP
renderer's reference-estimator mode; only ``tests/``, ``__graft_entry__``'s
smoke check and ``bench.py``'s CPU-baseline leg may import it.  The product
never calls it.

It restates ``ReferencePathTracer.render_hdr``
(``ordinarylight/renderers/reference/path_tracer.py:31-95``) operation for
operation, keeping every intermediate dtype of the reference (float32 state,
float64 primary directions, the mixed-precision Moller-Trumbore test and the
float32 ``best_t`` that makes the *last* of several near-equal hits in index
order win), because the renderer's parity is defined against that exact
sample stream:

* camera rays: ``path_tracer.py:109-151``;
* intersection: ``path_tracer.py:153-176``;
* geometric normals: ``path_tracer.py:178-182``;
* sky: ``path_tracer.py:184-202``;
* cosine-hemisphere continuation: ``path_tracer.py:204-211``;
* random stream: ``numpy.random.default_rng(seed)`` (PCG64), consumed in the
  order jitter-x, jitter-y, then one ``normal((n_hit, 3))`` per bounce with
  at least one hit (``path_tracer.py:39-87``).

Pinned by ``tests/golden/e0_reference_vectors.npz`` (outputs of the
reference functions themselves on small inputs) and by closed forms
(``tests/test_oracle_e0.py``).
"""

from __future__ import annotations

import math

import numpy as np


def camera_rays(camera, width, height, rng):
    """Primary rays: float32 origins, float64 directions (row-major pixels)."""
    position = np.asarray(camera.position, dtype=np.float32)
    forward = np.asarray(camera.target, dtype=np.float32) - position
    forward /= np.linalg.norm(forward)
    right = np.cross(forward, np.asarray(camera.up, dtype=np.float32))
    right /= np.linalg.norm(right)
    up = np.cross(right, forward)
    rows, cols = np.mgrid[0:height, 0:width]
    jx = rng.random((height, width), dtype=np.float32)
    jy = rng.random((height, width), dtype=np.float32)
    # int64 + float32 promotes to float64: the NDC grid is float64.
    ndc_x = (cols + jx) / width * 2.0 - 1.0
    ndc_y = 1.0 - (rows + jy) / height * 2.0
    aspect = width / height
    kind = type(camera).__name__
    if kind == "PanoramicCamera":
        yaw = ndc_x * math.radians(camera.horizontal_fov_degrees) * 0.5
        pitch = ndc_y * math.radians(camera.vertical_fov_degrees) * 0.5
        cp = np.cos(pitch)[..., None]
        dirs = (cp * np.cos(yaw)[..., None] * forward
                + cp * np.sin(yaw)[..., None] * right
                + np.sin(pitch)[..., None] * up).reshape((-1, 3))
        origins = np.repeat(position[None, :], len(dirs), axis=0)
    elif kind == "OrthographicCamera":
        half = camera.vertical_size * 0.5
        origins = (position[None, None, :]
                   + ndc_x[..., None] * (aspect * half) * right
                   + ndc_y[..., None] * half * up).reshape((-1, 3))
        dirs = np.repeat(forward[None, :], len(origins), axis=0)
    else:
        half = math.tan(math.radians(camera.vertical_fov_degrees) / 2.0)
        dirs = (forward[None, None, :]
                + ndc_x[..., None] * (aspect * half) * right
                + ndc_y[..., None] * half * up).reshape((-1, 3))
        origins = np.repeat(position[None, :], len(dirs), axis=0)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    return origins, dirs


def intersect(origins, directions, triangles):
    """Brute-force closest hit in triangle-index order.

    Returns ``(best_t float32, best_triangle int32)``; -1 marks a miss.
    """
    n = len(origins)
    best_t = np.full(n, np.inf, dtype=np.float32)
    best = np.full(n, -1, dtype=np.int32)
    for index, tri in enumerate(triangles):
        e1 = tri[1] - tri[0]
        e2 = tri[2] - tri[0]
        pvec = np.cross(directions, e2)
        det = pvec @ e1
        valid = np.abs(det) > 1e-7
        inv = np.zeros_like(det)
        inv[valid] = 1.0 / det[valid]
        tvec = origins - tri[0]
        u = np.sum(tvec * pvec, axis=1) * inv
        qvec = np.cross(tvec, e1)
        v = np.sum(directions * qvec, axis=1) * inv
        t = (qvec @ e2) * inv
        hit = valid & (u >= 0.0) & (v >= 0.0) & (u + v <= 1.0) & (t > 1e-4)
        closer = hit & (t < best_t)
        best_t[closer] = t[closer]
        best[closer] = index
    return best_t, best


def triangle_normals(triangles):
    """Unit geometric normals cross(b - a, c - a), float32."""
    n = np.cross(triangles[:, 1] - triangles[:, 0], triangles[:, 2] - triangles[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return n


def sky(directions, environment=None):
    """Miss radiance: environment colour x intensity (nearest texel) or gradient."""
    if environment is not None:
        radiance = np.ones((len(directions), 3), np.float32)
        if environment.image is not None:
            lon = np.arctan2(directions[:, 2], directions[:, 0])
            lon += environment.rotation
            u = np.mod(lon / (2.0 * np.pi) + 0.5, 1.0)
            v = np.arccos(np.clip(directions[:, 1], -1.0, 1.0)) / np.pi
            h, w, _ = environment.image.shape
            x = np.minimum((u * w).astype(np.int64), w - 1)
            y = np.minimum((v * h).astype(np.int64), h - 1)
            radiance = environment.image[y, x]
        return radiance * np.asarray(environment.color, np.float32) * float(environment.intensity)
    blend = np.clip(0.5 * (directions[:, 1] + 1.0), 0.0, 1.0)[:, None]
    return (1.0 - blend) * np.array((0.8, 0.8, 0.8)) + blend * np.array((0.2, 0.4, 0.9))


def cosine_directions(normals, gaussian):
    """Continuation directions from float32-cast normal deviates."""
    s = gaussian.astype(np.float32)
    s /= np.linalg.norm(s, axis=1, keepdims=True)
    s[np.sum(s * normals, axis=1) < 0.0] *= -1.0
    d = normals + s
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return d


def render_hdr(triangles, colors, emissions, camera, width, height, *,
               samples=1, max_bounces=3, seed=1, environment=None, trace=None):
    """Linear HDR float32 (H, W, 4), alpha 0.

    ``trace``, when given, is a list that receives one
    ``(sample, bounce, active_ray_ids, hit_triangle)`` tuple per bounce so
    primary-hit IDs and active sets can be compared.
    """
    if width <= 0 or height <= 0:
        raise ValueError("width and height must be positive")
    if samples <= 0 or max_bounces <= 0:
        raise ValueError("samples and max_bounces must be positive")
    triangles = np.asarray(triangles, np.float32)
    colors = np.asarray(colors, np.float32)
    emissions = np.asarray(emissions, np.float32)
    rng = np.random.default_rng(seed)
    total = np.zeros((height * width, 3), dtype=np.float32)
    for sample in range(samples):
        origins, dirs = camera_rays(camera, width, height, rng)
        radiance = np.zeros_like(origins)
        throughput = np.ones_like(origins)
        active = np.ones(len(origins), dtype=bool)
        for bounce in range(max_bounces):
            ids = np.flatnonzero(active)
            if not len(ids):
                break
            t, tri = intersect(origins[ids], dirs[ids], triangles)
            if trace is not None:
                trace.append((sample, bounce, ids.copy(), tri.copy()))
            missed = tri < 0
            if np.any(missed):
                m = ids[missed]
                radiance[m] += throughput[m] * sky(dirs[m], environment)
                active[m] = False
            hit = ~missed
            if not np.any(hit):
                continue
            h = ids[hit]
            tri_h = tri[hit]
            points = origins[h] + dirs[h] * t[hit, None]
            normals = triangle_normals(triangles[tri_h])
            normals[np.sum(normals * dirs[h], axis=1) > 0.0] *= -1.0
            radiance[h] += throughput[h] * emissions[tri_h]
            throughput[h] *= colors[tri_h]
            origins[h] = points + normals * 1e-4
            dirs[h] = cosine_directions(normals, rng.normal(size=normals.shape))
        total += radiance
    rgba = np.empty((height, width, 4), dtype=np.float32)
    rgba[..., :3] = total.reshape((height, width, 3)) / samples
    rgba[..., 3] = 0.0
    return rgba


def render_scene(scene, camera, width, height, *, samples=1, max_bounces=3,
                 seed=1, frame_index=0, trace=None):
    """``CpuReferenceRenderer.render_frame`` semantics (seed + frame_index)."""
    tris, colors, emissions = scene.triangles()
    return render_hdr(tris, colors, emissions, camera, width, height,
                      samples=samples, max_bounces=max_bounces,
                      seed=int(seed) + int(frame_index),
                      environment=scene.environment, trace=trace)
