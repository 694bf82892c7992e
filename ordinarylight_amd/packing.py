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

from . import textures as tx
from .scene import material_record

# (texture, transform) fields in the reference's binding order
# (Scene.textures / texture_bindings, scene/_core.py:2927-2976); the last four
# share the base-colour transform there.  Only the first six are sampled by
# the GI shading; the rest keep binding indices equal to the reference's.
_BINDING_SLOTS = (
    ("base_color_texture", "base_color_transform"),
    ("metallic_roughness_texture", "metallic_roughness_transform"),
    ("emissive_texture", "emissive_transform"),
    ("normal_texture", "normal_transform"),
    ("occlusion_texture", "occlusion_transform"),
    ("transmission_texture", "transmission_transform"),
    ("thickness_texture", "thickness_transform"),
    ("clearcoat_texture", "base_color_transform"),
    ("sheen_texture", "base_color_transform"),
    ("anisotropy_texture", "base_color_transform"),
    ("subsurface_texture", "base_color_transform"),
)
_TEXTURE_FIELDS = tuple(name for name, _ in _BINDING_SLOTS)
# Material record word (row 8 .w, after the sheen colour: a padding word of the reference record)
# that marks a material with texture bindings for the device.
TEXTURED_FLAG = (8, 3)


def _slot_pairs(material):
    for tex_name, tf_name in _BINDING_SLOTS:
        tex = getattr(material, tex_name, None)
        if tex is not None:
            yield tex, getattr(material, tf_name)


def is_textured(material):
    return any(True for _ in _slot_pairs(material))


class _Bindings:
    """Textures and (texture, transform) bindings in first-use order."""

    def __init__(self):
        self.textures, self.bindings = [], []

    def add(self, material):
        for tex, tf in _slot_pairs(material):
            if all(tex is not t for t in self.textures):
                self.textures.append(tex)
            if all(tex is not t or tf != f for t, f in self.bindings):
                self.bindings.append((tex, tf))

    def index(self, material, tex_name, tf_name):
        tex = getattr(material, tex_name, None)
        if tex is None:
            return -1.0
        tf = getattr(material, tf_name)
        return float(next(i for i, (t, f) in enumerate(self.bindings) if t is tex and f == tf))


def _material_block(material, bindings):
    """One material record with texture binding indices
    (Scene.triangle_material_data, scene/_core.py:2816-2924).  Untextured
    records hold the resolved occlusion factor 1 in row 5 .w (the value the
    reference's textured path computes when no occlusion map is bound);
    textured ones keep the transmission binding there, as the reference
    does, and set the flag."""
    block = material_record(material)[None]
    if not is_textured(material):
        block[:, 5, 3] = 1.0
        return block
    idx = lambda slot: bindings.index(material, *_BINDING_SLOTS[slot])  # noqa: E731
    uses_uv1 = any(tf.texcoord_set == 1 for _, tf in _slot_pairs(material))
    block[:, 3, 2] = block[:, 3, 2] + (0.25 if uses_uv1 else 0.0)
    block[:, 4] = (idx(0), idx(1), idx(2), idx(3))
    block[:, 5, 1] = idx(4)
    block[:, 5, 3] = idx(5)
    block[:, 10] = (idx(7), idx(8), idx(9), idx(10))
    block[:, 11, 0] = idx(6)
    block[:, TEXTURED_FLAG[0], TEXTURED_FLAG[1]] = 1.0
    return block


def _environment_words(environment, analytic, analytic_count, texture_list):
    """Texture words for ``texture_list`` plus an image environment's texture
    (appended last, as Scene.textures does); the environment record's row 3
    becomes (texture index, rotation, log range, 0) (analytic_light_data,
    scene/_core.py:2613-2629).  None when there is nothing to sample."""
    env_texture, log_range = tx.environment_texture(environment)
    if env_texture is not None:
        texture_list.append(env_texture)
        for k in range(min(analytic_count, len(analytic))):
            if int(analytic[k, 0, 3] + 0.5) == 3:
                analytic[k, 3, 0] = float(len(texture_list) - 1)
                analytic[k, 3, 2] = log_range
    return tx.texture_words(texture_list) if texture_list else None


def expanded_materials(packed):
    """(records [T, 12, 4], ids [T]): one full material record per triangle
    with its per-triangle parameters written in, the form the reference's
    Scene.triangle_material_data has; (materials, material_ids) unchanged
    when the scene has none."""
    if packed.face_params is None:
        return packed.materials, packed.material_ids
    records = packed.materials[packed.material_ids].copy()
    records[:, 0, :3] = packed.face_params[:, :3]
    records[:, 0, 3] = packed.face_params[:, 3]
    records[:, 1, 3] = packed.face_params[:, 4]
    return np.ascontiguousarray(records, np.float32), np.arange(len(records), dtype=np.uint32)


def _uv_density(world, tri, uv):
    """sqrt(uv area / world area) per triangle (triangle_attribute_data,
    scene/_core.py:2042-2069)."""
    p = world[tri]
    w = uv[tri]
    e = w[:, 1:] - w[:, :1]
    uv_area = np.abs(e[:, 0, 0] * e[:, 1, 1] - e[:, 0, 1] * e[:, 1, 0])
    we = p[:, 1:] - p[:, :1]
    world_area = np.linalg.norm(np.cross(we[:, 0], we[:, 1]), axis=1)
    return np.sqrt(uv_area / np.maximum(world_area, np.float32(1e-8))).astype(np.float32)


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
    # Per-triangle GGX parameters (meshes with face materials; None when no
    # mesh has them): base colour rgb, roughness, metallic replacing those
    # words of the triangle's material record; other triangles repeat their
    # record's values.  expanded_materials() gives one full record per
    # triangle instead.
    face_params: np.ndarray | None = None    # float32 [T, 5]
    # Textured scenes only (None otherwise):
    uvs: np.ndarray | None = None            # float32 [V, 4] uv0.xy, uv1.xy
    tangents: np.ndarray | None = None       # float32 [V, 4] world tangent, handedness
    uv_density: np.ndarray | None = None     # float32 [T] uv0 texels-per-world-unit scale
    texture_words: np.ndarray | None = None  # uint32 (textures.texture_words)
    bindings: np.ndarray | None = None       # float32 [B, 2, 4]

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
    meshes = [m for m in _meshes(scene) if len(m.indices)]
    bindings = _Bindings()
    for mesh in meshes:
        bindings.add(mesh.material)
    textured = bool(bindings.textures) or getattr(getattr(scene, "environment", None), "image", None) is not None
    uvs, tangents, density = [], [], []
    any_faces = any(getattr(m, "face_materials", None) is not None for m in meshes)
    face_params = []
    for mesh in meshes:
        tri = np.asarray(mesh.indices, dtype=np.uint32)
        world = np.asarray(mesh.world_vertices, dtype=np.float32)
        positions.append(world)
        normals.append(np.asarray(mesh.world_normals, dtype=np.float32))
        indices.append(tri + np.uint32(offset))
        offset += len(world)
        if bindings.textures:
            uv0 = np.asarray(mesh.texcoords, np.float32)
            uvs.append(np.column_stack((uv0, np.asarray(mesh.texcoords1, np.float32))))
            tangents.append(np.asarray(mesh.world_tangents, np.float32))
            density.append(_uv_density(world, tri, uv0))
        material = mesh.material
        key = id(material)
        if key not in material_index:
            material_index[key] = sum(len(m) for m in materials)
            materials.append(_material_block(material, bindings))
        material_ids.append(np.full(len(tri), material_index[key], np.uint32))
        if any_faces:
            faces = getattr(mesh, "face_materials", None)
            if faces is not None:
                face_params.append(np.column_stack((faces.base_color, faces.roughness, faces.metallic)))
            else:
                rec = materials[material_index[key]][0]
                face_params.append(np.broadcast_to(np.array((*rec[0], rec[1, 3]), np.float32), (len(tri), 5)))
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
    analytic = np.array(scene.analytic_light_data(), np.float32).reshape(-1, 4, 4)
    words = binding = None
    if textured:
        texture_list = list(bindings.textures)
        words = _environment_words(getattr(scene, "environment", None), analytic, analytic_count, texture_list)
        binding = tx.binding_records(bindings.bindings, texture_list)
    return PackedScene(
        positions=cat(positions, (0, 3), np.float32),
        normals=cat(normals, (0, 3), np.float32),
        indices=cat(indices, (0, 3), np.uint32),
        material_ids=cat(material_ids, (0,), np.uint32),
        materials=np.ascontiguousarray(np.concatenate(materials), np.float32),
        analytic_lights=np.ascontiguousarray(analytic, np.float32),
        analytic_light_count=analytic_count,
        area_lights=np.ascontiguousarray(scene.emissive_triangle_data(), np.float32),
        area_light_count=area_count,
        area_light_weight=float(scene.emissive_light_weight),
        environment=getattr(scene, "environment", None),
        has_textures=textured,
        face_params=cat(face_params, (0, 5), np.float32) if any_faces else None,
        uvs=cat(uvs, (0, 4), np.float32) if bindings.textures else None,
        tangents=cat(tangents, (0, 4), np.float32) if bindings.textures else None,
        uv_density=cat(density, (0,), np.float32) if bindings.textures else None,
        texture_words=words,
        bindings=binding,
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
    has_textures: bool                       # some material has texture bindings
    texture_words: np.ndarray | None = None  # an image environment's texture

    @property
    def triangle_count(self):
        return int(self.instance_triangle_offset[-1] + (self.triangle_offsets[self.instance_resource[-1] + 1]
                                                        - self.triangle_offsets[self.instance_resource[-1]]))


def has_textured_materials(scene):
    """True when some rendered mesh's material has texture bindings."""
    return any(is_textured(m.material) for m in _meshes(scene) if len(m.indices))


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
            if is_textured(material):
                raise ValueError("materials with texture bindings need a flat scene (instanced=False)")
            record = material_record(material)
            record[5, 3] = 1.0  # resolved occlusion factor of an untextured material
            materials.append(record[None])
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
    analytic = np.array(scene.analytic_light_data(), np.float32).reshape(-1, 4, 4)
    words = _environment_words(getattr(scene, "environment", None), analytic, analytic_count, [])
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
        analytic_lights=np.ascontiguousarray(analytic, np.float32),
        analytic_light_count=analytic_count,
        area_lights=np.ascontiguousarray(scene.emissive_triangle_data(), np.float32),
        area_light_count=area_count,
        area_light_weight=float(scene.emissive_light_weight),
        environment=getattr(scene, "environment", None),
        has_textures=textured,
        texture_words=words,
    )


__all__ = ["PackedScene", "PackedInstancedScene", "expanded_materials", "has_textured_materials", "pack_scene",
           "pack_instanced_scene", "shares_resources"]
