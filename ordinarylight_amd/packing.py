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
        key = id(material)
        if key not in material_index:
            material_index[key] = len(materials)
            if any(getattr(material, name, None) is not None for name in _TEXTURE_FIELDS):
                textured = True
            record = material_record(material)
            # Resolved occlusion factor of an untextured material: the value
            # the reference's textured path computes when no occlusion map is
            # bound (mix(1, 1, strength) = 1).
            record[5, 3] = 1.0
            materials.append(record)
        material_ids.append(np.full(len(tri), material_index[key], np.uint32))
    if not materials:
        record = material_record(type("Default", (), {
            "base_color": (0.8, 0.8, 0.8), "emission": (0.0, 0.0, 0.0),
            "roughness": 1.0, "metallic": 0.0})())
        record[5, 3] = 1.0
        materials.append(record)
    cat = lambda parts, shape, dtype: (np.ascontiguousarray(np.concatenate(parts), dtype)  # noqa: E731
                                       if parts else np.zeros(shape, dtype))
    analytic_count = int(getattr(scene, "analytic_light_count", 0))
    area_count = int(getattr(scene, "emissive_triangle_count", 0))
    return PackedScene(
        positions=cat(positions, (0, 3), np.float32),
        normals=cat(normals, (0, 3), np.float32),
        indices=cat(indices, (0, 3), np.uint32),
        material_ids=cat(material_ids, (0,), np.uint32),
        materials=np.ascontiguousarray(np.stack(materials), np.float32),
        analytic_lights=np.ascontiguousarray(scene.analytic_light_data(), np.float32),
        analytic_light_count=analytic_count,
        area_lights=np.ascontiguousarray(scene.emissive_triangle_data(), np.float32),
        area_light_count=area_count,
        area_light_weight=float(scene.emissive_light_weight),
        environment=getattr(scene, "environment", None),
        has_textures=textured,
    )


__all__ = ["PackedScene", "pack_scene"]
