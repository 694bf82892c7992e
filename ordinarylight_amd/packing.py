"""Scene adapter: any OrdinaryLight-style scene -> the olhip device format.

Only duck-typed attributes are read (``render_meshes``/``visible_meshes``,
``world_vertices``, ``world_normals``, ``indices``, ``material``,
``analytic_light_data()``, ``emissive_triangle_data()``,
``emissive_light_weight``, ``environment``), so both this package's
:class:`~ordinarylight_amd.scene.Scene` and a reference ``Scene`` are
accepted.

Device format (indexed, compact) instead of the reference's flattened
per-triangle buffers (``targets/vulkan/scene.py:1345-1380``: 48 + 192 + 144
bytes per triangle): world-space vertices and unit normals per vertex, three
indices and a material index per triangle (16 bytes), and one 12 x vec4
material record per distinct material.  Triangle order is the reference's
packed order (render meshes in insertion order), so triangle IDs agree with
``Scene.triangle_instance_ids()`` and ``triangle_material_ids()``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .scene import material_record

_TEXTURE_FIELDS = (
    "base_color_texture", "metallic_roughness_texture", "emissive_texture",
    "normal_texture", "occlusion_texture", "transmission_texture",
    "thickness_texture", "clearcoat_texture", "sheen_texture",
    "anisotropy_texture", "subsurface_texture",
)


@dataclass
class PackedScene:
    positions: np.ndarray       # float32 [V, 3]
    normals: np.ndarray         # float32 [V, 3]
    indices: np.ndarray         # uint32 [T, 3]
    material_ids: np.ndarray    # uint32 [T]
    materials: np.ndarray       # float32 [M, 12, 4]
    analytic_lights: np.ndarray  # float32 [P, 4, 4]
    analytic_light_count: int
    area_lights: np.ndarray     # float32 [L, 5, 4]
    area_light_count: int
    area_light_weight: float
    environment: object
    has_textures: bool

    @property
    def triangle_count(self):
        return int(len(self.indices))


def _meshes(scene):
    meshes = getattr(scene, "render_meshes", None)
    if meshes is None:
        meshes = scene.visible_meshes
    return tuple(meshes)


def pack_scene(scene):
    """Pack ``scene`` into flat arrays for ``olhip_scene_create``."""
    positions, normals, indices, material_ids = [], [], [], []
    materials, material_index = [], {}
    offset = 0
    textured = False
    for mesh in _meshes(scene):
        tri = np.asarray(mesh.indices, dtype=np.uint32)
        if not len(tri):
            continue
        world = np.asarray(mesh.world_vertices, dtype=np.float32)
        positions.append(world)
        normals.append(np.asarray(mesh.world_normals, dtype=np.float32))
        indices.append(tri + np.uint32(offset))
        offset += len(world)
        material = mesh.material
        if any(getattr(material, name, None) is not None for name in _TEXTURE_FIELDS):
            textured = True
        faces = getattr(mesh, "face_materials", None)
        if faces is not None:
            # One record per triangle, appended in triangle order.
            block = faces.records(material)
            block[:, 5, 3] = 1.0
            first = sum(len(m) for m in materials)
            materials.append(block)
            material_ids.append(np.arange(first, first + len(tri), dtype=np.uint32))
            continue
        key = id(material)
        if key not in material_index:
            material_index[key] = sum(len(m) for m in materials)
            record = material_record(material)
            # Resolved occlusion factor of an untextured material: the value
            # the reference's textured path computes when no occlusion map is
            # bound (mix(1, 1, strength) = 1).
            record[5, 3] = 1.0
            materials.append(record[None])
        material_ids.append(np.full(len(tri), material_index[key], np.uint32))
    if not materials:
        record = material_record(type("Default", (), {
            "base_color": (0.8, 0.8, 0.8), "emission": (0.0, 0.0, 0.0),
            "roughness": 1.0, "metallic": 0.0})())
        record[5, 3] = 1.0
        materials.append(record[None])
    cat = lambda parts, shape, dtype: (np.ascontiguousarray(np.concatenate(parts), dtype)  # noqa: E731
                                       if parts else np.zeros(shape, dtype))
    analytic_count = int(getattr(scene, "analytic_light_count", 0))
    area_count = int(getattr(scene, "emissive_triangle_count", 0))
    return PackedScene(
        positions=cat(positions, (0, 3), np.float32),
        normals=cat(normals, (0, 3), np.float32),
        indices=cat(indices, (0, 3), np.uint32),
        material_ids=cat(material_ids, (0,), np.uint32),
        materials=np.ascontiguousarray(np.concatenate(materials), np.float32),
        analytic_lights=np.ascontiguousarray(scene.analytic_light_data(), np.float32),
        analytic_light_count=analytic_count,
        area_lights=np.ascontiguousarray(scene.emissive_triangle_data(), np.float32),
        area_light_count=area_count,
        area_light_weight=float(scene.emissive_light_weight),
        environment=getattr(scene, "environment", None),
        has_textures=textured,
    )


@dataclass
class PackedInstancedScene:
    """Two-level layout: object-space resources and their placements."""

    vertex_offsets: np.ndarray      # uint32 [R + 1]
    triangle_offsets: np.ndarray    # uint32 [R + 1]
    positions: np.ndarray           # float32 [V, 3] object space
    normals: np.ndarray             # float32 [V, 3]
    indices: np.ndarray             # uint32 [T, 3] local to each resource
    instance_resource: np.ndarray   # uint32 [I]
    instance_material: np.ndarray   # uint32 [I]
    instance_triangle_offset: np.ndarray  # uint32 [I] first triangle in packed order
    instance_id: np.ndarray         # uint32 [I]
    world_to_object: np.ndarray     # float32 [I, 3, 4]
    normal_matrix: np.ndarray       # float32 [I, 3, 3]
    determinant: np.ndarray         # float32 [I]
    world_bounds: np.ndarray        # float32 [I, 6]
    world_spheres: np.ndarray       # float32 [I, 4] centre xyz, radius (contains the instance)
    materials: np.ndarray
    analytic_lights: np.ndarray
    analytic_light_count: int
    area_lights: np.ndarray
    area_light_count: int
    area_light_weight: float
    environment: object
    has_textures: bool

    @property
    def triangle_count(self):
        return int(self.instance_triangle_offset[-1] + (self.triangle_offsets[self.instance_resource[-1] + 1]
                                                        - self.triangle_offsets[self.instance_resource[-1]]))


def shares_resources(scene):
    """True when some mesh resource is placed by more than one visible mesh."""
    seen = set()
    for mesh in _meshes(scene):
        resource = getattr(mesh, "resource", None)
        if resource is None or not len(mesh.indices):
            continue
        if id(resource) in seen:
            return True
        seen.add(id(resource))
    return False


def pack_instanced_scene(scene):
    """Pack ``scene`` as resources + instances for ``olhip_scene_create_instanced``.

    Triangle order stays the flattened packed order (instance k's triangles
    start at ``instance_triangle_offset[k]``), so ids and tie-breaks agree
    with :func:`pack_scene` and ``Scene.triangle_instance_ids()``.
    """
    meshes = [m for m in _meshes(scene) if len(m.indices)]
    if not meshes:
        raise ValueError("a two-level scene needs at least one mesh")
    resources, res_index = [], {}
    materials, material_index = [], {}
    textured = False
    inst_res, inst_mat, inst_off, inst_id, mats = [], [], [], [], []
    offset = 0
    for mesh in meshes:
        if getattr(mesh, "face_materials", None) is not None:
            raise ValueError("per-triangle materials need a flat scene (instanced=False)")
        resource = mesh.resource
        if id(resource) not in res_index:
            res_index[id(resource)] = len(resources)
            resources.append(resource)
        material = mesh.material
        if id(material) not in material_index:
            material_index[id(material)] = len(materials)
            if any(getattr(material, name, None) is not None for name in _TEXTURE_FIELDS):
                textured = True
            record = material_record(material)
            record[5, 3] = 1.0  # resolved occlusion factor of an untextured material
            materials.append(record)
        inst_res.append(res_index[id(resource)])
        inst_mat.append(material_index[id(material)])
        inst_off.append(offset)
        inst_id.append(int(mesh.id) if mesh.id is not None else len(inst_id))
        mats.append(np.asarray(mesh.transform.matrix, np.float64))
        offset += len(mesh.indices)
    vcount = np.array([len(r.vertices) for r in resources], np.int64)
    tcount = np.array([len(r.indices) for r in resources], np.int64)
    vertex_offsets = np.concatenate(([0], np.cumsum(vcount))).astype(np.uint32)
    triangle_offsets = np.concatenate(([0], np.cumsum(tcount))).astype(np.uint32)
    positions = np.ascontiguousarray(np.concatenate([r.vertices for r in resources]), np.float32)
    normals = np.ascontiguousarray(np.concatenate([r.normals for r in resources]), np.float32)
    indices = np.ascontiguousarray(np.concatenate([r.indices for r in resources]), np.uint32)
    M = np.stack(mats)                        # [I, 4, 4] object -> world
    inv = np.linalg.inv(M)                    # world -> object
    A = M[:, :3, :3]
    normal_matrix = np.transpose(np.linalg.inv(A), (0, 2, 1))
    det = np.linalg.det(A)
    # World bounds, in float64 and rounded outward to float32: the box of the
    # 8 transformed corners of each resource's object box, intersected with
    # the box of its transformed bounding sphere (centre c, radius r: world
    # axis i spans (A c + t)_i +- r |A_i|).  Both contain the instance; for
    # rotated round meshes the sphere box is up to sqrt(3) narrower per axis.
    lo = np.stack([r.vertices.min(0) for r in resources]).astype(np.float64)
    hi = np.stack([r.vertices.max(0) for r in resources]).astype(np.float64)
    centre = 0.5 * (lo + hi)
    radius = np.array([np.sqrt(((r.vertices.astype(np.float64) - c) ** 2).sum(1).max())
                       for r, c in zip(resources, centre)])
    ri = np.asarray(inst_res)
    corners = np.stack([np.where(np.array(bits, bool), hi[ri], lo[ri])
                        for bits in np.ndindex(2, 2, 2)], axis=1)      # [I, 8, 3]
    world = corners @ np.transpose(A, (0, 2, 1)) + M[:, None, :3, 3]
    sphere_c = np.einsum("kij,kj->ki", A, centre[ri]) + M[:, :3, 3]
    sphere_h = radius[ri][:, None] * np.linalg.norm(A, axis=2)
    bmin = np.maximum(world.min(1), sphere_c - sphere_h)
    bmax = np.minimum(world.max(1), sphere_c + sphere_h)
    wmin = np.nextafter(bmin.astype(np.float32), np.float32(-np.inf))
    wmax = np.nextafter(bmax.astype(np.float32), np.float32(np.inf))
    # World bounding spheres: the transformed centre, rounded to float32, and
    # the radius scaled by the largest singular value of the linear part, grown
    # by the centre's rounding and rounded up.
    centre32 = sphere_c.astype(np.float32)
    grow = np.linalg.norm(centre32.astype(np.float64) - sphere_c, axis=1)
    reach = radius[ri] * np.linalg.norm(A, ord=2, axis=(1, 2)) * (1.0 + 1e-6) + grow
    reach32 = np.nextafter(reach.astype(np.float32), np.float32(np.inf))
    spheres = np.concatenate((centre32, reach32[:, None]), 1)
    if not materials:
        raise ValueError("no materials")
    analytic_count = int(getattr(scene, "analytic_light_count", 0))
    area_count = int(getattr(scene, "emissive_triangle_count", 0))
    return PackedInstancedScene(
        vertex_offsets=vertex_offsets, triangle_offsets=triangle_offsets,
        positions=positions, normals=normals, indices=indices,
        instance_resource=np.asarray(inst_res, np.uint32), instance_material=np.asarray(inst_mat, np.uint32),
        instance_triangle_offset=np.asarray(inst_off, np.uint32), instance_id=np.asarray(inst_id, np.uint32),
        world_to_object=np.ascontiguousarray(inv[:, :3, :4], np.float32),
        normal_matrix=np.ascontiguousarray(normal_matrix, np.float32),
        determinant=np.ascontiguousarray(det, np.float32),
        world_bounds=np.ascontiguousarray(np.concatenate((wmin, wmax), 1), np.float32),
        world_spheres=np.ascontiguousarray(spheres, np.float32),
        materials=np.ascontiguousarray(np.concatenate(materials), np.float32),
        analytic_lights=np.ascontiguousarray(scene.analytic_light_data(), np.float32),
        analytic_light_count=analytic_count,
        area_lights=np.ascontiguousarray(scene.emissive_triangle_data(), np.float32),
        area_light_count=area_count,
        area_light_weight=float(scene.emissive_light_weight),
        environment=getattr(scene, "environment", None),
        has_textures=textured,
    )


__all__ = ["PackedScene", "PackedInstancedScene", "pack_scene", "pack_instanced_scene", "shares_resources"]
