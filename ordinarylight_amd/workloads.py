"""Seeded procedural scenes for the renderer's benchmark and test workloads.

The reference ships no room, terrain or icosphere generators; these are the
workload definitions of this package, all deterministic in their seed:

* :func:`box_room` -- closed box room of 36 triangles: 5 walls (10), two
  boxes with a seeded rotation about +Y (24), one emissive ceiling quad (2)
  of radiance 15.  Albedos are uniform in [0.1, 0.9]^3.  Pinhole camera.
* :func:`terrain` -- fBm-displaced heightfield of ``(nx-1)*(nz-1)*2``
  triangles (1001 x 501 vertices give 1,000,000) in chunk meshes, with a
  GGX material drawn per triangle (albedo ~ U[0.1, 0.9]^3, roughness ~
  U[0.05, 1], metallic ~ Bernoulli(0.3)), plus 64 emissive triangles and a
  constant environment of radiance 0.5.
* :func:`tessellated_room`, :func:`mixed_scene`, :func:`mixed_jobs` -- the
  batched/served workload: eight scenes of 10k to 2M triangles from the two
  generators and a seeded list of jobs over them.
"""

from __future__ import annotations

import math

import numpy as np

from .cameras import PerspectiveCamera
from .lights import EnvironmentLight
from .scene import FaceMaterials, Material, Scene, Texture, TextureTransform, Transform


def _quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3); normal = (p1-p0) x (p2-p0)."""
    vertices = np.asarray((p0, p1, p2, p3), np.float32)
    indices = np.asarray(((0, 1, 2), (0, 2, 3)), np.uint32)
    return vertices, indices


def _box(size, albedo_rng):
    """Axis-aligned box centred on the origin's footprint, 12 outward triangles."""
    sx, sy, sz = (float(v) * 0.5 for v in size)
    corners = np.asarray([
        (-sx, 0.0, -sz), (sx, 0.0, -sz), (sx, 0.0, sz), (-sx, 0.0, sz),
        (-sx, 2 * sy, -sz), (sx, 2 * sy, -sz), (sx, 2 * sy, sz), (-sx, 2 * sy, sz),
    ], np.float32)
    faces = (
        (0, 1, 2, 3),  # bottom (normal -y)
        (4, 7, 6, 5),  # top (+y)
        (0, 4, 5, 1),  # back (-z)
        (3, 2, 6, 7),  # front (+z)
        (0, 3, 7, 4),  # left (-x)
        (1, 5, 6, 2),  # right (+x)
    )
    indices = []
    for a, b, c, d in faces:
        indices.extend(((a, b, c), (a, c, d)))
    return corners, np.asarray(indices, np.uint32)


def box_room(seed=0):
    """Closed box room (36 triangles) and its pinhole camera."""
    rng = np.random.default_rng(seed)
    albedo = lambda: tuple(float(v) for v in rng.uniform(0.1, 0.9, 3))  # noqa: E731
    scene = Scene()
    # Room spans x in [-1, 1], y in [0, 2], z in [-1, 1]; walls face inward.
    walls = (
        ((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1)),    # floor (+y)
        ((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)),    # ceiling (-y)
        ((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1)),  # back (+z)
        ((-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (-1, 0, 1)),  # left (+x)
        ((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1)),      # right (-x)
    )
    for corners in walls:
        vertices, indices = _quad(*corners)
        scene.add_mesh(vertices, indices, Material(base_color=albedo()))
    for size, position in (((0.55, 1.2, 0.55), (-0.38, 0.0, -0.3)),
                           ((0.55, 0.6, 0.55), (0.4, 0.0, 0.35))):
        vertices, indices = _box(size, rng)
        angle = float(rng.uniform(-0.6, 0.6))
        transform = Transform.translation(position) @ Transform.rotation((0, 1, 0), angle)
        scene.add_mesh(vertices, indices, Material(base_color=albedo()), transform=transform)
    light_v, light_i = _quad((-0.25, 1.99, -0.25), (0.25, 1.99, -0.25),
                             (0.25, 1.99, 0.25), (-0.25, 1.99, 0.25))
    scene.add_mesh(light_v, light_i, Material(base_color=(0.8, 0.8, 0.8),
                                              emission=(15.0, 15.0, 15.0)))
    camera = PerspectiveCamera(position=(0.0, 1.0, 3.4), target=(0.0, 1.0, 0.0),
                               vertical_fov_degrees=40.0)
    return scene, camera


def _texture(rng, h, w, kind, **sampler):
    """Seeded RGBA8 test textures: checker colours, random bytes, or a
    normal map of random tilts (tangent-space, z up)."""
    if kind == "checker":
        a, b = rng.integers(20, 236, 3), rng.integers(20, 236, 3)
        yy, xx = np.mgrid[0:h, 0:w]
        px = np.where(((yy // 4 + xx // 4) % 2)[..., None] == 0, a, b)
        px = np.concatenate((px, np.full((h, w, 1), 255)), axis=2)
    elif kind == "normal":
        v = rng.normal(0.0, 0.35, (h, w, 3))
        v[..., 2] = 1.0
        v /= np.linalg.norm(v, axis=2, keepdims=True)
        px = np.concatenate((np.round((v * 0.5 + 0.5) * 255), np.full((h, w, 1), 255)), axis=2)
    else:
        px = rng.integers(0, 256, (h, w, 4))
    return Texture(np.clip(px, 0, 255).astype(np.uint8), **sampler)


def _planar_uv(vertices, scale):
    """Texture coordinates by planar projection (x + z, y), scaled."""
    v = np.asarray(vertices, np.float32)
    return np.ascontiguousarray(np.column_stack((v[:, 0] + v[:, 2], v[:, 1])) * np.float32(scale), np.float32)


def textured_room(seed=0, open_sky=False):
    """The box room with textured materials: base colour, metallic-roughness,
    emissive, normal, occlusion and transmission maps; repeat, clamp and
    mirror wrapping, nearest and bilinear filtering, transformed bindings and
    a second texture-coordinate set.  ``open_sky`` keeps the floor and the
    boxes under an image environment instead of the walls and ceiling."""
    rng = np.random.default_rng(seed)
    scene = Scene()
    floor_v, floor_i = _quad((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1))
    floor = Material(
        base_color=(0.9, 0.85, 0.8), roughness=0.8, metallic=0.6,
        base_color_texture=_texture(rng, 32, 32, "checker"),
        base_color_transform=TextureTransform(offset=(0.1, 0.2), scale=(3.0, 2.5), rotation=0.3),
        metallic_roughness_texture=_texture(rng, 16, 16, "random"),
        normal_texture=_texture(rng, 24, 24, "normal"), normal_scale=0.8)
    uv = _planar_uv(floor_v[:, [0, 2, 1]], 0.9)
    scene.add_mesh(floor_v, floor_i, floor, texcoords=uv, texcoords1=uv[:, ::-1] * 1.7)
    if not open_sky:
        back_v, back_i = _quad((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1))
        back = Material(
            base_color=(0.8, 0.8, 0.8), roughness=0.5,
            base_color_texture=_texture(rng, 16, 24, "checker", wrap_s="mirror", wrap_t="repeat"),
            base_color_transform=TextureTransform(scale=(2.2, 1.5), texcoord_set=1),
            normal_texture=_texture(rng, 16, 16, "normal"), normal_transform=TextureTransform(texcoord_set=1))
        uv = _planar_uv(back_v, 1.0)
        scene.add_mesh(back_v, back_i, back, texcoords=uv, texcoords1=uv * np.float32(1.3) + np.float32(0.2))
        for corners, texture in (
                (((-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (-1, 0, 1)),
                 _texture(rng, 8, 8, "checker", wrap_s="clamp", wrap_t="clamp", linear_filter=False)),
                (((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1)), _texture(rng, 12, 20, "checker", wrap_t="mirror")),
                (((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)), None)):
            v, i = _quad(*corners)
            mat = Material(base_color=tuple(float(c) for c in rng.uniform(0.3, 0.9, 3)),
                           base_color_texture=texture,
                           base_color_transform=TextureTransform(offset=(-0.3, 0.15), scale=(1.6, 1.4)))
            scene.add_mesh(v, i, mat, texcoords=_planar_uv(v, 1.1))
        light_v, light_i = _quad((-0.25, 1.99, -0.25), (0.25, 1.99, -0.25), (0.25, 1.99, 0.25), (-0.25, 1.99, 0.25))
        scene.add_mesh(light_v, light_i, Material(base_color=(0.8, 0.8, 0.8), emission=(12.0, 12.0, 12.0),
                                                  emissive_texture=_texture(rng, 8, 8, "checker")),
                       texcoords=_planar_uv(light_v[:, [0, 2, 1]], 2.0))
    box_v, box_i = _box((0.55, 1.2, 0.55), rng)
    occl = Material(base_color=(0.7, 0.6, 0.5), roughness=0.4,
                    base_color_texture=_texture(rng, 16, 16, "checker"),
                    occlusion_texture=_texture(rng, 8, 8, "random"), occlusion_strength=0.7)
    scene.add_mesh(box_v, box_i, occl, texcoords=_planar_uv(box_v, 1.5),
                   transform=Transform.translation((-0.38, 0.0, -0.3)) @ Transform.rotation((0, 1, 0), 0.4))
    box_v, box_i = _box((0.55, 0.6, 0.55), rng)
    glass = Material(base_color=(0.9, 0.95, 1.0), transmission=1.0, ior=1.5,
                     transmission_texture=_texture(rng, 8, 8, "random"),
                     base_color_texture=_texture(rng, 8, 8, "checker"))
    scene.add_mesh(box_v, box_i, glass, texcoords=_planar_uv(box_v, 1.2),
                   transform=Transform.translation((0.4, 0.0, 0.35)) @ Transform.rotation((0, 1, 0), -0.3))
    if open_sky:
        yy, xx = np.mgrid[0:16, 0:32]
        sky = np.stack((0.4 + 0.3 * np.sin(xx / 5.0), 0.5 + 0.2 * np.cos(yy / 3.0), 0.9 + 0.0 * xx), axis=2)
        sky[2:5, 6:9] = (40.0, 36.0, 30.0)  # a bright sun
        scene.set_environment(EnvironmentLight(image=sky.astype(np.float32), color=(1.0, 0.95, 0.9),
                                               intensity=0.8, rotation=0.7))
    camera = PerspectiveCamera(position=(0.0, 1.0, 3.4), target=(0.0, 0.7, 0.0), vertical_fov_degrees=42.0)
    return scene, camera


def _value_noise(rng, shape_cells):
    """Lattice of uniform values for bilinear value noise."""
    return rng.uniform(-1.0, 1.0, size=(shape_cells[0] + 2, shape_cells[1] + 2))


def _fbm(x, z, rng, octaves=6):
    """Fractional Brownian motion of smoothstep-interpolated value noise."""
    height = np.zeros_like(x, dtype=np.float64)
    amplitude, frequency = 1.0, 1.0 / 96.0
    for _ in range(octaves):
        cells = (int(np.ceil(x.max() * frequency)) + 1, int(np.ceil(z.max() * frequency)) + 1)
        lattice = _value_noise(rng, cells)
        fx, fz = x * frequency, z * frequency
        ix, iz = np.floor(fx).astype(np.int64), np.floor(fz).astype(np.int64)
        tx, tz = fx - ix, fz - iz
        tx, tz = tx * tx * (3 - 2 * tx), tz * tz * (3 - 2 * tz)
        v00, v10 = lattice[ix, iz], lattice[ix + 1, iz]
        v01, v11 = lattice[ix, iz + 1], lattice[ix + 1, iz + 1]
        height += amplitude * ((v00 * (1 - tx) + v10 * tx) * (1 - tz)
                               + (v01 * (1 - tx) + v11 * tx) * tz)
        amplitude *= 0.5
        frequency *= 2.0
    return height


def terrain(seed=0, nx=1001, nz=501, chunk=50, emissive_quads=32, height_scale=60.0, materials="triangle"):
    """fBm heightfield terrain with GGX materials, emitters and sky.

    ``materials="triangle"`` draws every triangle's base colour (U[0.1,0.9]^3),
    roughness (U[0.05,1]) and metallic factor (Bernoulli(0.3)) on its own;
    ``"chunk"`` draws one material per chunk mesh."""
    if materials not in ("triangle", "chunk"):
        raise ValueError("materials must be 'triangle' or 'chunk'")
    rng = np.random.default_rng(seed)
    xs = np.arange(nx, dtype=np.float64)
    zs = np.arange(nz, dtype=np.float64)
    gx, gz = np.meshgrid(xs, zs, indexing="ij")
    heights = _fbm(gx, gz, rng) * height_scale
    positions = np.stack((gx, heights, gz), axis=-1).astype(np.float32)  # (nx, nz, 3)
    # Area-weighted vertex normals over the whole grid (continuous at chunk seams).
    flat = positions.reshape(-1, 3)
    vid = np.arange(nx * nz, dtype=np.int64).reshape(nx, nz)
    a, b = vid[:-1, :-1].ravel(), vid[1:, :-1].ravel()
    c, d = vid[1:, 1:].ravel(), vid[:-1, 1:].ravel()
    # (a, d, c) and (a, c, b) wind so the face normal points to +y.
    tri_all = np.concatenate((np.stack((a, d, c), 1), np.stack((a, c, b), 1)))
    face = np.cross(flat[tri_all[:, 1]] - flat[tri_all[:, 0]], flat[tri_all[:, 2]] - flat[tri_all[:, 0]])
    normals = np.zeros_like(flat)
    for corner in range(3):
        np.add.at(normals, tri_all[:, corner], face)
    normals /= np.maximum(np.linalg.norm(normals, axis=1, keepdims=True), 1e-12)
    normals = normals.reshape(nx, nz, 3)

    scene = Scene()
    n_tri = (nx - 1) * (nz - 1) * 2
    if materials == "triangle":
        face_base = rng.uniform(0.1, 0.9, (n_tri, 3)).astype(np.float32)
        face_rough = rng.uniform(0.05, 1.0, n_tri).astype(np.float32)
        face_metal = (rng.random(n_tri) < 0.3).astype(np.float32)
    first = 0
    for x0 in range(0, nx - 1, chunk):
        for z0 in range(0, nz - 1, chunk):
            x1, z1 = min(x0 + chunk, nx - 1), min(z0 + chunk, nz - 1)
            cx, cz = x1 - x0 + 1, z1 - z0 + 1
            verts = positions[x0:x1 + 1, z0:z1 + 1].reshape(-1, 3)
            norms = normals[x0:x1 + 1, z0:z1 + 1].reshape(-1, 3)
            lid = np.arange(cx * cz, dtype=np.uint32).reshape(cx, cz)
            la, lb = lid[:-1, :-1].ravel(), lid[1:, :-1].ravel()
            lc, ld = lid[1:, 1:].ravel(), lid[:-1, 1:].ravel()
            quads = np.stack((la, ld, lc, lb), 1)
            indices = np.empty((len(quads) * 2, 3), np.uint32)
            indices[0::2] = quads[:, (0, 1, 2)]
            indices[1::2] = quads[:, (0, 2, 3)]
            if materials == "triangle":
                rows = slice(first, first + len(indices))
                first += len(indices)
                scene.add_mesh(verts, indices, Material(), normals=norms,
                               face_materials=FaceMaterials(face_base[rows], face_rough[rows], face_metal[rows]))
                continue
            material = Material(
                base_color=tuple(float(v) for v in rng.uniform(0.1, 0.9, 3)),
                roughness=float(rng.uniform(0.05, 1.0)),
                metallic=float(rng.random() < 0.3),
            )
            scene.add_mesh(verts, indices, material, normals=norms)
    # 64 emissive triangles: 32 small quads hovering above the surface, facing down.
    lv, li = [], []
    for q in range(emissive_quads):
        px = float(rng.uniform(0.05, 0.95) * (nx - 1))
        pz = float(rng.uniform(0.05, 0.95) * (nz - 1))
        ground = heights[int(px), int(pz)]
        py = float(ground + rng.uniform(8.0, 20.0))
        s = float(rng.uniform(2.0, 5.0))
        base = len(lv)
        lv.extend(((px - s, py, pz - s), (px + s, py, pz - s), (px + s, py, pz + s), (px - s, py, pz + s)))
        li.extend(((base, base + 1, base + 2), (base, base + 2, base + 3)))
    scene.add_mesh(np.asarray(lv, np.float32), np.asarray(li, np.uint32),
                   Material(base_color=(0.9, 0.9, 0.9), emission=(12.0, 10.5, 8.0)))
    scene.set_environment(EnvironmentLight(color=(1.0, 1.0, 1.0), intensity=0.5))
    mid_height = float(np.median(heights))
    camera = PerspectiveCamera(
        position=(0.5 * (nx - 1), mid_height + 0.5 * height_scale + 40.0, -0.35 * (nz - 1)),
        target=(0.5 * (nx - 1), mid_height, 0.55 * (nz - 1)),
        vertical_fov_degrees=50.0,
    )
    return scene, camera


def _tessellate(corners, n):
    """Planar quad (p0, p1, p2, p3) as an n x n grid: 2 n^2 triangles wound like _quad."""
    p0, p1, p2, p3 = (np.asarray(c, np.float64) for c in corners)
    t = np.linspace(0.0, 1.0, n + 1)
    u, v = np.meshgrid(t, t, indexing="ij")
    u, v = u[..., None], v[..., None]
    grid = (1 - u) * (1 - v) * p0 + u * (1 - v) * p1 + u * v * p2 + (1 - u) * v * p3
    vid = np.arange((n + 1) * (n + 1), dtype=np.uint32).reshape(n + 1, n + 1)
    a, b = vid[:-1, :-1].ravel(), vid[1:, :-1].ravel()
    c, d = vid[1:, 1:].ravel(), vid[:-1, 1:].ravel()
    indices = np.empty((2 * n * n, 3), np.uint32)
    indices[0::2] = np.stack((a, b, c), 1)
    indices[1::2] = np.stack((a, c, d), 1)
    return grid.reshape(-1, 3).astype(np.float32), indices


def tessellated_room(seed=0, target_triangles=10_000):
    """:func:`box_room` with its five walls tessellated into n x n grids so the
    room holds about ``target_triangles`` (same materials, boxes, light and
    camera as the 36-triangle room of the same seed)."""
    rng = np.random.default_rng(seed)
    albedo = lambda: tuple(float(v) for v in rng.uniform(0.1, 0.9, 3))  # noqa: E731
    n = max(1, int(round(math.sqrt(max(target_triangles - 26, 10) / 10.0))))
    scene = Scene()
    walls = (
        ((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1)),
        ((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)),
        ((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1)),
        ((-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (-1, 0, 1)),
        ((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1)),
    )
    for corners in walls:
        vertices, indices = _tessellate(corners, n)
        scene.add_mesh(vertices, indices, Material(base_color=albedo()))
    for size, position in (((0.55, 1.2, 0.55), (-0.38, 0.0, -0.3)),
                           ((0.55, 0.6, 0.55), (0.4, 0.0, 0.35))):
        vertices, indices = _box(size, rng)
        angle = float(rng.uniform(-0.6, 0.6))
        transform = Transform.translation(position) @ Transform.rotation((0, 1, 0), angle)
        scene.add_mesh(vertices, indices, Material(base_color=albedo()), transform=transform)
    light_v, light_i = _quad((-0.25, 1.99, -0.25), (0.25, 1.99, -0.25),
                             (0.25, 1.99, 0.25), (-0.25, 1.99, 0.25))
    scene.add_mesh(light_v, light_i, Material(base_color=(0.8, 0.8, 0.8),
                                              emission=(15.0, 15.0, 15.0)))
    camera = PerspectiveCamera(position=(0.0, 1.0, 3.4), target=(0.0, 1.0, 0.0),
                               vertical_fov_degrees=40.0)
    return scene, camera


def icosphere(subdivisions=5, seed=0, jitter=0.02):
    """Unit icosphere with 20 * 4^subdivisions triangles; vertices moved
    radially by a seeded relative jitter.  Returns (vertices, indices)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    v = np.asarray(verts, np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.asarray(faces, np.int64)
    for _ in range(subdivisions):
        edges = np.concatenate((f[:, (0, 1)], f[:, (1, 2)], f[:, (2, 0)]))
        key = np.sort(edges, axis=1)
        uniq, inverse = np.unique(key, axis=0, return_inverse=True)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = inverse.reshape(3, -1).T + len(v)  # midpoints of (01, 12, 20)
        v = np.concatenate((v, mid))
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        f = np.concatenate((np.stack((a, m[:, 0], m[:, 2]), 1), np.stack((b, m[:, 1], m[:, 0]), 1),
                            np.stack((c, m[:, 2], m[:, 1]), 1), np.stack((m[:, 0], m[:, 1], m[:, 2]), 1)))
    rng = np.random.default_rng(seed)
    v *= 1.0 + jitter * rng.uniform(-1.0, 1.0, (len(v), 1))
    return v.astype(np.float32), f.astype(np.uint32)


def _random_rotation(rng):
    """Uniform random rotation (Shoemake's quaternion method), 3x3."""
    u1, u2, u3 = rng.random(3)
    q = (math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
         math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3))
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def instanced_spheres(seed=0, meshes=5, subdivisions=5, instances=10_000, box=120.0, dielectric=0.2,
                      emitters=100_000, emitter_size=0.35):
    """Two-level scene: ``meshes`` jittered icospheres (20 * 4^subdivisions
    triangles each) placed ``instances`` times under random affine
    transforms (uniform rotation, scale in [0.5, 2], translation in a cube of
    side ``box``); a ``dielectric`` fraction of the instances is glass
    (transmission 1, IOR 1.5), the rest GGX; ``emitters`` small emissive
    triangles scattered through the cube."""
    rng = np.random.default_rng(seed)
    scene = Scene()
    resources = []
    for k in range(meshes):
        v, f = icosphere(subdivisions, seed=seed * 1000 + k)
        resources.append(scene.create_mesh(v, f, Material()))
    glass = Material(base_color=(0.95, 0.97, 1.0), transmission=1.0, ior=1.5, roughness=0.0)
    palette = [Material(base_color=tuple(float(x) for x in rng.uniform(0.1, 0.9, 3)),
                        roughness=float(rng.uniform(0.05, 1.0)), metallic=float(rng.random() < 0.3))
               for _ in range(64)]
    half = 0.5 * box
    for _ in range(instances):
        m = np.eye(4)
        m[:3, :3] = _random_rotation(rng) * rng.uniform(0.5, 2.0)
        m[:3, 3] = rng.uniform(-half, half, 3)
        material = glass if rng.random() < dielectric else palette[int(rng.integers(len(palette)))]
        scene.add_instance(resources[int(rng.integers(meshes))], transform=Transform(m), material=material)
    if emitters:
        centers = rng.uniform(-half, half, (emitters, 1, 3))
        offsets = rng.normal(0.0, emitter_size, (emitters, 3, 3))
        ev = (centers + offsets).reshape(-1, 3).astype(np.float32)
        ei = np.arange(3 * emitters, dtype=np.uint32).reshape(-1, 3)
        scene.add_mesh(ev, ei, Material(base_color=(0.9, 0.9, 0.9), emission=(14.0, 12.0, 9.0),
                                        emission_two_sided=True))
    camera = PerspectiveCamera(position=(0.0, 0.0, -0.75 * box), target=(0.0, 0.0, 0.0),
                               vertical_fov_degrees=55.0)
    return scene, camera


# Resident scenes of the batched and served workloads: three rooms and five
# terrains spanning 10k to 2M triangles (terrain with (2k+1) x (k+1) vertices
# has 4 k^2 triangles).
MIXED_SCENES = (
    ("room_10k", "room", 10_000),
    ("room_40k", "room", 40_000),
    ("room_160k", "room", 160_000),
    ("terrain_100k", "terrain", 158),
    ("terrain_250k", "terrain", 250),
    ("terrain_500k", "terrain", 354),
    ("terrain_1m", "terrain", 500),
    ("terrain_2m", "terrain", 707),
)


def mixed_scene(key, seed=0):
    """One scene of :data:`MIXED_SCENES` by key."""
    for name, kind, size in MIXED_SCENES:
        if name == key:
            if kind == "room":
                return tessellated_room(seed, size)
            return terrain(seed, nx=2 * size + 1, nz=size + 1, chunk=50, emissive_quads=32)
    raise KeyError(key)


def mixed_jobs(seed=0, n_jobs=64, scene_keys=None, min_edge=128, max_edge=1024):
    """Job list of the batched workload: square images whose edge is
    log-uniform in [min_edge, max_edge], a scene drawn uniformly from
    ``scene_keys`` (default: all of :data:`MIXED_SCENES`), samples per pixel
    uniform in {1, 2, 4, 8, 16} and max_bounces in {4, 8}."""
    rng = np.random.default_rng(seed)
    keys = list(scene_keys) if scene_keys is not None else [name for name, _, _ in MIXED_SCENES]
    jobs = []
    for _ in range(n_jobs):
        edge = int(round(math.exp(rng.uniform(math.log(min_edge), math.log(max_edge)))))
        jobs.append({
            "scene": keys[int(rng.integers(len(keys)))],
            "width": edge,
            "height": edge,
            "samples": int(rng.choice((1, 2, 4, 8, 16))),
            "max_bounces": int(rng.choice((4, 8))),
        })
    return jobs


__all__ = ["box_room", "terrain", "tessellated_room", "MIXED_SCENES", "mixed_scene", "mixed_jobs", "icosphere",
           "instanced_spheres"]
