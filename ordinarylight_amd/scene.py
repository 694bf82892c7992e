"""Scene model: meshes, reusable mesh resources, instances, materials.

This is the API subset of OrdinaryLight's scene model that the HIP renderer
consumes (``ordinarylight/scene/_core.py``).  Field names, defaults and the
packed layouts returned by the ``*_data`` methods follow the reference so
that a reference ``Scene`` and this one are interchangeable for the
renderer's scene adapter (:mod:`ordinarylight_amd.packing`):

* ``Material`` fields and defaults: ``scene/_core.py:431-474``;
* ``Mesh`` normal generation (area-weighted, welded by position):
  ``scene/_core.py:716-738``;
* ``Scene.triangles()``: ``scene/_core.py:2732-2753``;
* ``emissive_triangle_data`` / ``emissive_light_weight`` /
  ``emissive_triangle_count``: ``scene/_core.py:2637-2730``;
* ``analytic_light_data``: ``scene/_core.py:2578-2634``;
* ``triangle_instance_ids`` / ``triangle_material_ids``:
  ``scene/_core.py:2773-2805``.

Unlike the reference, no per-triangle Python loops run here, so packing a
million-triangle mesh takes well under a second.
"""

from __future__ import annotations

from dataclasses import dataclass, field
import math
from types import MappingProxyType

import numpy as np

from .lights import (
    DirectionalLight, EnvironmentLight, LIGHT_TYPES, PointLight, SpotLight,
)

# Light type codes of the packed analytic-light records
# (``ordinarylight/scene/_core.py:2578-2634``; environment is type 3).
POINT, DIRECTIONAL, SPOT, ENVIRONMENT = 0, 1, 2, 3

_LUMINANCE = (0.2126, 0.7152, 0.0722)


def _vec3(value, name):
    array = np.asarray(value, dtype=np.float32)
    if array.shape != (3,) or not np.all(np.isfinite(array)):
        raise ValueError(f"{name} must contain three finite values")
    return array


@dataclass(frozen=True)
class Transform:
    """Immutable affine object-to-world transform (row-major 4x4)."""

    matrix: np.ndarray = field(
        default_factory=lambda: np.eye(4, dtype=np.float32),
        compare=False, repr=False,
    )

    def __post_init__(self):
        matrix = np.array(self.matrix, dtype=np.float32, copy=True)
        if matrix.shape != (4, 4):
            raise ValueError("transform matrix must have shape (4, 4)")
        if not np.all(np.isfinite(matrix)):
            raise ValueError("transform matrix must contain finite values")
        if not np.allclose(matrix[3], (0.0, 0.0, 0.0, 1.0), atol=1e-6):
            raise ValueError("transform matrix must be affine")
        if abs(float(np.linalg.det(matrix[:3, :3]))) < 1e-10:
            raise ValueError("transform linear component must be invertible")
        matrix.setflags(write=False)
        object.__setattr__(self, "matrix", matrix)

    @classmethod
    def translation(cls, offset):
        matrix = np.eye(4, dtype=np.float32)
        matrix[:3, 3] = _vec3(offset, "translation")
        return cls(matrix)

    @classmethod
    def scale(cls, factors):
        values = np.asarray(factors, dtype=np.float32)
        if values.ndim == 0:
            values = np.repeat(values, 3)
        matrix = np.eye(4, dtype=np.float32)
        matrix[:3, :3] = np.diag(_vec3(values, "scale"))
        return cls(matrix)

    @classmethod
    def rotation(cls, axis, angle_radians):
        """Rodrigues rotation about ``axis`` (right-handed)."""
        axis = _vec3(axis, "rotation axis").astype(np.float64)
        length = float(np.linalg.norm(axis))
        if length < 1e-8:
            raise ValueError("rotation axis cannot be zero")
        x, y, z = axis / length
        c, s = math.cos(angle_radians), math.sin(angle_radians)
        k = np.array(((0.0, -z, y), (z, 0.0, -x), (-y, x, 0.0)))
        rotation = c * np.eye(3) + s * k + (1.0 - c) * np.outer((x, y, z), (x, y, z))
        matrix = np.eye(4, dtype=np.float32)
        matrix[:3, :3] = rotation
        return cls(matrix)

    def __matmul__(self, other):
        if not isinstance(other, Transform):
            return NotImplemented
        return Transform(self.matrix @ other.matrix)


_WRAP_MODES = {"repeat": 0, "clamp": 1, "mirror": 2}


@dataclass(frozen=True)
class Texture:
    """RGBA8 image with glTF sampler state (the reference's ``Texture``,
    ``scene/_core.py:116-135``): wrap per axis and bilinear/nearest filter.
    Textures are bound by identity, as the reference binds them."""

    pixels: np.ndarray = field(compare=False, repr=False)
    wrap_s: str = "repeat"
    wrap_t: str = "repeat"
    linear_filter: bool = True

    def __post_init__(self):
        pixels = np.ascontiguousarray(self.pixels, dtype=np.uint8)
        if pixels.ndim != 3 or pixels.shape[2] != 4:
            raise ValueError("texture pixels must have shape (height, width, 4)")
        if pixels.shape[0] < 1 or pixels.shape[1] < 1:
            raise ValueError("texture dimensions must be positive")
        for name in ("wrap_s", "wrap_t"):
            if getattr(self, name) not in _WRAP_MODES:
                raise ValueError(f"{name} must be repeat, clamp, or mirror")
        if not isinstance(self.linear_filter, bool):
            raise TypeError("linear_filter must be a bool")
        pixels.flags.writeable = False
        object.__setattr__(self, "pixels", pixels)


@dataclass(frozen=True)
class TextureTransform:
    """Affine UV mapping of one binding (KHR_texture_transform, as the
    reference's ``TextureTransform``, ``scene/_core.py:403-425``): scale, then
    rotate, then offset, on texture coordinate set 0 or 1."""

    offset: tuple = (0.0, 0.0)
    scale: tuple = (1.0, 1.0)
    rotation: float = 0.0
    texcoord_set: int = 0

    def __post_init__(self):
        if len(self.offset) != 2 or len(self.scale) != 2:
            raise ValueError("texture transform offset and scale must be vec2 values")
        offset = tuple(float(v) for v in self.offset)
        scale = tuple(float(v) for v in self.scale)
        rotation = float(self.rotation)
        if int(self.texcoord_set) not in (0, 1):
            raise ValueError("texcoord_set must be zero or one")
        if not all(math.isfinite(v) for v in (*offset, *scale, rotation)):
            raise ValueError("texture transform values must be finite")
        if abs(scale[0]) < 1e-8 or abs(scale[1]) < 1e-8:
            raise ValueError("texture transform scale components cannot be zero")
        object.__setattr__(self, "offset", offset)
        object.__setattr__(self, "scale", scale)
        object.__setattr__(self, "rotation", rotation)
        object.__setattr__(self, "texcoord_set", int(self.texcoord_set))


# Texture slots the GI shading applies (shadeApplyMaterialTextures and
# shadeApplyNormalTexture, wavefront_shade_candidate.glsl:915-955), each with
# its transform field.
MATERIAL_TEXTURE_SLOTS = (
    ("base_color_texture", "base_color_transform"),
    ("metallic_roughness_texture", "metallic_roughness_transform"),
    ("emissive_texture", "emissive_transform"),
    ("normal_texture", "normal_transform"),
    ("occlusion_texture", "occlusion_transform"),
    ("transmission_texture", "transmission_transform"),
)


@dataclass(frozen=True)
class Material:
    """Physically based material; defaults follow the reference material."""

    base_color: tuple = (0.8, 0.8, 0.8)
    emission: tuple = (0.0, 0.0, 0.0)
    metallic: float = 0.0
    roughness: float = 1.0
    transmission: float = 0.0
    clearcoat: float = 0.0
    clearcoat_roughness: float = 0.1
    sheen_color: tuple = (0.0, 0.0, 0.0)
    sheen_roughness: float = 0.5
    anisotropy: float = 0.0
    thin_walled: bool = False
    subsurface: float = 0.0
    subsurface_color: tuple = (1.0, 1.0, 1.0)
    subsurface_radius: float = 0.5
    ior: float = 1.5
    attenuation_color: tuple = (1.0, 1.0, 1.0)
    attenuation_distance: float = float("inf")
    thickness: float = 0.0
    opacity: float = 1.0
    alpha_mode: str = "opaque"
    alpha_cutoff: float = 0.5
    emission_two_sided: bool = False
    normal_scale: float = 1.0
    occlusion_strength: float = 1.0
    base_color_texture: Texture | None = None
    base_color_transform: TextureTransform = TextureTransform()
    metallic_roughness_texture: Texture | None = None
    metallic_roughness_transform: TextureTransform = TextureTransform()
    emissive_texture: Texture | None = None
    emissive_transform: TextureTransform = TextureTransform()
    normal_texture: Texture | None = None
    normal_transform: TextureTransform = TextureTransform()
    occlusion_texture: Texture | None = None
    occlusion_transform: TextureTransform = TextureTransform()
    transmission_texture: Texture | None = None
    transmission_transform: TextureTransform = TextureTransform()

    def __post_init__(self):
        for texture_name, transform_name in MATERIAL_TEXTURE_SLOTS:
            if getattr(self, texture_name) is not None and not isinstance(getattr(self, texture_name), Texture):
                raise TypeError(f"{texture_name} must be a Texture or None")
            if not isinstance(getattr(self, transform_name), TextureTransform):
                raise TypeError(f"{transform_name} must be a TextureTransform")
        color = _vec3(self.base_color, "base_color")
        emission = _vec3(self.emission, "emission")
        attenuation = _vec3(self.attenuation_color, "attenuation_color")
        _vec3(self.sheen_color, "sheen_color")
        _vec3(self.subsurface_color, "subsurface_color")
        if np.any(color < 0.0) or np.any(color > 1.0):
            raise ValueError("base_color components must be between zero and one")
        if np.any(emission < 0.0):
            raise ValueError("emission components cannot be negative")
        for name in ("metallic", "roughness", "transmission", "clearcoat",
                     "clearcoat_roughness", "sheen_roughness", "subsurface",
                     "subsurface_radius", "opacity", "alpha_cutoff",
                     "occlusion_strength"):
            if not 0.0 <= float(getattr(self, name)) <= 1.0:
                raise ValueError(f"{name} must be between zero and one")
        if not -1.0 <= self.anisotropy <= 1.0:
            raise ValueError("anisotropy must be between minus one and one")
        if self.ior <= 0.0:
            raise ValueError("ior must be positive")
        if np.any(attenuation < 0.0) or np.any(attenuation > 1.0):
            raise ValueError("attenuation_color components must be between zero and one")
        if self.attenuation_distance <= 0.0:
            raise ValueError("attenuation_distance must be positive")
        if self.alpha_mode not in {"opaque", "mask", "blend"}:
            raise ValueError("alpha_mode must be opaque, mask, or blend")
        if not isinstance(self.thin_walled, bool):
            raise TypeError("thin_walled must be a bool")
        if not isinstance(self.emission_two_sided, bool):
            raise TypeError("emission_two_sided must be a bool")


def material_record(material, program_id=0):
    """Pack one material as the reference's 12 x vec4 record.

    Layout of ``Scene.triangle_material_data`` (``scene/_core.py:2836-2923``)
    with every texture binding set to -1 (untextured).
    """
    g = lambda name, default: getattr(material, name, default)  # noqa: E731
    return np.asarray((
        (*material.base_color, material.roughness),
        (*material.emission, material.metallic),
        (*g("attenuation_color", (1.0, 1.0, 1.0)), g("transmission", 0.0)),
        (g("ior", 1.5), g("attenuation_distance", float("inf")),
         float(program_id), float(g("emission_two_sided", False))),
        (-1.0, -1.0, -1.0, -1.0),
        (g("normal_scale", 1.0), -1.0, g("occlusion_strength", 1.0), -1.0),
        (g("clearcoat", 0.0), g("clearcoat_roughness", 0.1),
         g("sheen_roughness", 0.5), g("anisotropy", 0.0)),
        (g("subsurface", 0.0), g("subsurface_radius", 0.5),
         float(g("thin_walled", False)), g("thickness", 0.0)),
        (*g("sheen_color", (0.0, 0.0, 0.0)), 0.0),
        (*g("subsurface_color", (1.0, 1.0, 1.0)), 0.0),
        (-1.0, -1.0, -1.0, -1.0),
        (-1.0, g("opacity", 1.0), g("alpha_cutoff", 0.5),
         float({"opaque": 0, "mask": 1, "blend": 2}[g("alpha_mode", "opaque")])),
    ), dtype=np.float32)


@dataclass(frozen=True, eq=False)
class FaceMaterials:
    """Per-triangle GGX parameters of one mesh (an extension of the
    reference's one-material-per-mesh model, for workloads that draw a
    material per triangle): triangle k is shaded with the mesh material's
    record in which base colour, roughness and metallic factor are replaced
    by row k.  Emission and every other field stay the mesh material's."""

    base_color: np.ndarray   # float32 [T, 3] in [0, 1]
    roughness: np.ndarray    # float32 [T] in [0, 1]
    metallic: np.ndarray     # float32 [T] in [0, 1]

    def __post_init__(self):
        base = np.ascontiguousarray(self.base_color, np.float32)
        rough = np.ascontiguousarray(self.roughness, np.float32).reshape(-1)
        metal = np.ascontiguousarray(self.metallic, np.float32).reshape(-1)
        if base.ndim != 2 or base.shape[1] != 3 or not len(base) == len(rough) == len(metal):
            raise ValueError("face materials need base_color [T, 3], roughness [T] and metallic [T]")
        for name, a in (("base_color", base), ("roughness", rough), ("metallic", metal)):
            if not np.all(np.isfinite(a)) or a.min(initial=0.0) < 0.0 or a.max(initial=0.0) > 1.0:
                raise ValueError(f"face {name} values must lie in [0, 1]")
        object.__setattr__(self, "base_color", base)
        object.__setattr__(self, "roughness", rough)
        object.__setattr__(self, "metallic", metal)

    def __len__(self):
        return len(self.roughness)

    def records(self, material):
        """float32 [T, 12, 4] material records over ``material``'s record."""
        out = np.repeat(material_record(material)[None], len(self), axis=0)
        out[:, 0, :3] = self.base_color
        out[:, 0, 3] = self.roughness
        out[:, 1, 3] = self.metallic
        return out


def generate_normals(vertices, indices):
    """Area-weighted vertex normals, welded by identical position.

    Same definition as ``Mesh._generate_normals`` (``scene/_core.py:716-738``):
    unnormalised face normals summed into each corner, then summed over all
    vertices sharing a position, normalised, ``(0,1,0)`` where degenerate.
    """
    vertices = np.asarray(vertices, np.float32)
    indices = np.asarray(indices, np.uint32)
    normals = np.zeros_like(vertices)
    if len(indices):
        tri = vertices[indices]
        face = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
        for corner in range(3):
            np.add.at(normals, indices[:, corner], face)
    _, groups = np.unique(vertices, axis=0, return_inverse=True)
    groups = groups.reshape(-1)
    welded = np.zeros((int(groups.max()) + 1 if len(groups) else 0, 3), np.float32)
    np.add.at(welded, groups, normals)
    normals = welded[groups]
    lengths = np.linalg.norm(normals, axis=1)
    missing = lengths < 1e-8
    normals[~missing] /= lengths[~missing, None]
    normals[missing] = (0.0, 1.0, 0.0)
    return np.ascontiguousarray(normals, dtype=np.float32)


def generate_tangents(vertices, indices, normals, texcoords):
    """Per-vertex tangents (xyz, handedness) from texture-coordinate
    gradients, the definition of ``Mesh._generate_tangents``
    (``scene/_core.py:740-772``): per triangle the uv-gradient tangent and
    bitangent (triangles with |det| < 1e-10 skipped), summed into each
    corner, Gram-Schmidt against the normal (a fallback axis where the sum
    vanishes), handedness from the bitangent sum."""
    vertices = np.asarray(vertices, np.float32)
    indices = np.asarray(indices, np.int64)
    normals = np.asarray(normals, np.float32)
    uv = np.asarray(texcoords, np.float32)
    t_sum = np.zeros_like(vertices)
    b_sum = np.zeros_like(vertices)
    if len(indices):
        p = vertices[indices]
        w = uv[indices]
        ea, eb = p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]
        ua, ub = w[:, 1] - w[:, 0], w[:, 2] - w[:, 0]
        det = ua[:, 0] * ub[:, 1] - ua[:, 1] * ub[:, 0]
        ok = np.abs(det.astype(np.float64)) >= 1e-10
        inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0).astype(np.float32)
        tan = (ea * ub[:, 1:2] - eb * ua[:, 1:2]) * inv[:, None]
        bit = (eb * ua[:, 0:1] - ea * ub[:, 0:1]) * inv[:, None]
        for corner in range(3):
            np.add.at(t_sum, indices[ok, corner], tan[ok])
            np.add.at(b_sum, indices[ok, corner], bit[ok])
    tangent = t_sum - normals * np.sum(normals * t_sum, axis=1, keepdims=True)
    length = np.linalg.norm(tangent, axis=1)
    small = length < 1e-8
    if np.any(small):
        n = normals[small]
        axis = np.where(np.abs(n[:, 2:3]) < 0.999, np.array([[0.0, 0.0, 1.0]], np.float32),
                        np.array([[0.0, 1.0, 0.0]], np.float32))
        tangent[small] = np.cross(n, axis)
        length[small] = np.linalg.norm(tangent[small], axis=1)
    tangent = tangent / np.maximum(length, 1e-8)[:, None]
    hand = np.where(np.sum(np.cross(normals, tangent) * b_sum, axis=1) < 0.0, -1.0, 1.0)
    return np.ascontiguousarray(np.column_stack((tangent, hand)), dtype=np.float32)


@dataclass(eq=False)
class MeshResource:
    """Reusable object-space geometry shared by any number of instances.

    ``texcoords`` / ``texcoords1`` default to zeros and ``tangents`` are
    generated from ``texcoords`` when not given (``scene/_core.py:561-637``).
    """

    vertices: np.ndarray
    indices: np.ndarray
    material: Material = field(default_factory=Material)
    normals: np.ndarray | None = None
    name: str | None = None
    metadata: dict = field(default_factory=dict, repr=False)
    texcoords: np.ndarray | None = field(default=None, repr=False)
    texcoords1: np.ndarray | None = field(default=None, repr=False)
    tangents: np.ndarray | None = field(default=None, repr=False)
    id: int | None = field(default=None, init=False)

    def __post_init__(self):
        self.vertices = np.ascontiguousarray(self.vertices, dtype=np.float32)
        self.indices = np.ascontiguousarray(self.indices, dtype=np.uint32)
        if self.vertices.ndim != 2 or self.vertices.shape[1] != 3:
            raise ValueError("vertices must have shape (vertex_count, 3)")
        if self.indices.ndim != 2 or self.indices.shape[1] != 3:
            raise ValueError("indices must have shape (triangle_count, 3)")
        if self.indices.size and int(self.indices.max()) >= len(self.vertices):
            raise ValueError("an index refers to a vertex that does not exist")
        if self.normals is None:
            self.normals = generate_normals(self.vertices, self.indices)
        else:
            normals = np.ascontiguousarray(self.normals, dtype=np.float32)
            if normals.shape != self.vertices.shape:
                raise ValueError("normals must have the same shape as vertices")
            lengths = np.linalg.norm(normals, axis=1)
            if np.any(lengths < 1e-8):
                raise ValueError("normals cannot contain zero vectors")
            self.normals = normals / lengths[:, None]
        count = len(self.vertices)
        for name in ("texcoords", "texcoords1"):
            value = getattr(self, name)
            value = np.zeros((count, 2), np.float32) if value is None else np.ascontiguousarray(value, np.float32)
            if value.shape != (count, 2):
                raise ValueError(f"{name} must have shape (vertex_count, 2)")
            setattr(self, name, value)
        if self.tangents is not None:
            tangents = np.array(self.tangents, dtype=np.float32)
            if tangents.shape != (count, 4):
                raise ValueError("tangents must have shape (vertex_count, 4)")
            lengths = np.linalg.norm(tangents[:, :3], axis=1)
            if np.any(lengths < 1e-8):
                raise ValueError("tangent directions cannot contain zero vectors")
            tangents[:, :3] /= lengths[:, None]
            tangents[:, 3] = np.where(tangents[:, 3] < 0.0, -1.0, 1.0)
            self.tangents = np.ascontiguousarray(tangents)
        self.metadata = MappingProxyType(dict(self.metadata))

    def tangent_frame(self):
        """The given tangents, or ones generated from ``texcoords`` on first
        use (untextured scenes never need them) and kept in ``tangents``."""
        if self.tangents is None:
            self.tangents = generate_tangents(self.vertices, self.indices, self.normals, self.texcoords)
        return self.tangents


@dataclass(eq=False)
class Mesh:
    """A placed, shaded occurrence of a :class:`MeshResource`."""

    resource: MeshResource
    material: Material = field(default_factory=Material)
    transform: Transform = field(default_factory=Transform)
    visible: bool = True
    name: str | None = None
    metadata: dict = field(default_factory=dict, repr=False)
    face_materials: FaceMaterials | None = field(default=None, repr=False)
    id: int | None = field(default=None, init=False)

    def __post_init__(self):
        if not isinstance(self.transform, Transform):
            self.transform = Transform(self.transform)
        if self.face_materials is not None and len(self.face_materials) != len(self.resource.indices):
            raise ValueError("face_materials must hold one row per triangle of the mesh")
        if not isinstance(self.visible, bool):
            raise TypeError("visible must be a bool")
        self.metadata = MappingProxyType(dict(self.metadata))

    @property
    def vertices(self):
        return self.resource.vertices

    @property
    def indices(self):
        return self.resource.indices

    @property
    def normals(self):
        return self.resource.normals

    @property
    def world_vertices(self):
        """Object vertices transformed into world space (float32)."""
        linear = self.transform.matrix[:3, :3]
        translation = self.transform.matrix[:3, 3]
        return np.ascontiguousarray(
            self.vertices @ linear.T + translation, dtype=np.float32,
        )

    @property
    def world_normals(self):
        """Normals transformed by the inverse transpose, renormalised."""
        normal_matrix = np.linalg.inv(self.transform.matrix[:3, :3]).T
        result = self.normals @ normal_matrix.T
        result /= np.linalg.norm(result, axis=1, keepdims=True)
        return np.ascontiguousarray(result, dtype=np.float32)

    @property
    def texcoords(self):
        return self.resource.texcoords

    @property
    def texcoords1(self):
        return self.resource.texcoords1

    @property
    def tangents(self):
        return self.resource.tangent_frame()

    @property
    def world_tangents(self):
        """Tangents through the linear part, made orthogonal to the world
        normal and renormalised; handedness flips with a mirroring transform
        (``scene/_core.py:701-714``)."""
        linear = self.transform.matrix[:3, :3]
        directions = self.tangents[:, :3] @ linear.T
        normals = self.world_normals
        directions -= normals * np.sum(directions * normals, axis=1)[:, None]
        directions /= np.linalg.norm(directions, axis=1, keepdims=True)
        handedness = self.tangents[:, 3].copy()
        if np.linalg.det(linear) < 0.0:
            handedness *= -1.0
        return np.ascontiguousarray(np.column_stack((directions, handedness)), dtype=np.float32)


Instance = Mesh


@dataclass(eq=False)
class Scene:
    """Ordered collection of placed meshes, analytic lights and environment."""

    meshes: list = field(default_factory=list)
    lights: list = field(default_factory=list)
    mesh_resources: list = field(default_factory=list)
    environment: EnvironmentLight | None = None
    metadata: dict = field(default_factory=dict)
    _revision: int = field(default=0, init=False, repr=False)
    _next_id: int = field(default=1, init=False, repr=False)
    _next_resource_id: int = field(default=1, init=False, repr=False)
    _next_material_id: int = field(default=1, init=False, repr=False)
    _material_ids: dict = field(default_factory=dict, init=False, repr=False)

    def __post_init__(self):
        meshes, lights, resources = list(self.meshes), list(self.lights), list(self.mesh_resources)
        self.meshes, self.lights, self.mesh_resources = [], [], []
        for resource in resources:
            self._attach_resource(resource)
        for mesh in meshes:
            if mesh.resource not in self.mesh_resources:
                self._attach_resource(mesh.resource)
            self._attach_mesh(mesh)
        for light in lights:
            self.add_light(light)
        if self.environment is not None and not isinstance(self.environment, EnvironmentLight):
            raise TypeError("environment must be an EnvironmentLight or None")

    # -- bookkeeping -----------------------------------------------------
    @property
    def revision(self):
        return self._revision

    def _changed(self):
        self._revision += 1

    def _attach_resource(self, resource):
        if resource.id is not None:
            raise ValueError("a mesh resource cannot belong to multiple scenes")
        resource.id = self._next_resource_id
        self._next_resource_id += 1
        self.mesh_resources.append(resource)

    def _attach_mesh(self, mesh):
        if mesh.id is not None:
            raise ValueError("a scene resource cannot belong to multiple scenes")
        mesh.id = self._next_id
        self._next_id += 1
        self._register_material(mesh.material)
        self.meshes.append(mesh)
        self._changed()
        return mesh

    def _register_material(self, material):
        existing = self._material_ids.get(id(material))
        if existing is not None and existing[0] is material:
            return existing[1]
        material_id = self._next_material_id
        self._next_material_id += 1
        self._material_ids[id(material)] = (material, material_id)
        return material_id

    def material_id(self, material):
        """Stable scene-local ID of an attached material (first-use order)."""
        existing = self._material_ids.get(id(material))
        if existing is None or existing[0] is not material:
            raise KeyError("material is not attached to this scene")
        return existing[1]

    # -- construction ----------------------------------------------------
    def create_mesh(self, vertices, indices, material=None, *, normals=None,
                    name=None, metadata=None, texcoords=None, texcoords1=None, tangents=None):
        """Create reusable object-space geometry without placing it."""
        resource = MeshResource(vertices, indices, material or Material(),
                                normals=normals, name=name,
                                metadata=metadata or {}, texcoords=texcoords,
                                texcoords1=texcoords1, tangents=tangents)
        self._attach_resource(resource)
        return resource

    create_mesh_resource = create_mesh

    def add_instance(self, resource, *, transform=None, material=None,
                     visible=True, name=None, metadata=None, face_materials=None):
        """Place one independently shaded occurrence of ``resource``."""
        if resource not in self.mesh_resources:
            raise KeyError("mesh resource is not part of this scene")
        mesh = Mesh(resource, resource.material if material is None else material,
                    Transform() if transform is None else transform,
                    visible=visible, name=name, metadata=metadata or {},
                    face_materials=face_materials)
        return self._attach_mesh(mesh)

    def add_instances(self, resource, transforms, *, materials=None, visible=True):
        """Place many occurrences from column-oriented inputs."""
        transforms = [t if isinstance(t, Transform) else Transform(t) for t in transforms]
        if materials is None or isinstance(materials, Material):
            materials = [materials] * len(transforms)
        if len(materials) != len(transforms):
            raise ValueError("expected one material per transform")
        return tuple(
            self.add_instance(resource, transform=t, material=m, visible=visible)
            for t, m in zip(transforms, materials)
        )

    def add_mesh(self, vertices, indices, material=None, *, normals=None,
                 transform=None, name=None, metadata=None, face_materials=None,
                 texcoords=None, texcoords1=None, tangents=None):
        resource = self.create_mesh(vertices, indices, material,
                                    normals=normals, name=name, metadata=metadata,
                                    texcoords=texcoords, texcoords1=texcoords1, tangents=tangents)
        return self.add_instance(resource, transform=transform, name=name,
                                 metadata=metadata, face_materials=face_materials)

    def add_light(self, light):
        if not isinstance(light, LIGHT_TYPES):
            raise TypeError("light must be a PointLight, DirectionalLight, or SpotLight")
        if light.id is not None:
            raise ValueError("a light cannot belong to multiple scenes")
        light.id = self._next_id
        self._next_id += 1
        self.lights.append(light)
        self._changed()
        return light

    def add_point_light(self, position, color=(1.0, 1.0, 1.0), intensity=1.0, *, range=None):
        return self.add_light(PointLight(position, color, intensity, range))

    def add_directional_light(self, direction, color=(1.0, 1.0, 1.0), intensity=1.0):
        return self.add_light(DirectionalLight(direction, color, intensity))

    def add_spot_light(self, position, direction, color=(1.0, 1.0, 1.0), intensity=1.0,
                       *, range=None, inner_cone_angle=0.3, outer_cone_angle=0.5):
        return self.add_light(SpotLight(position, direction, color, intensity, range,
                                        inner_cone_angle, outer_cone_angle))

    def set_environment(self, environment=None, **parameters):
        if environment is not None and parameters:
            raise TypeError("pass an EnvironmentLight or keyword parameters, not both")
        if parameters:
            environment = EnvironmentLight(**parameters)
        if environment is not None and not isinstance(environment, EnvironmentLight):
            raise TypeError("environment must be an EnvironmentLight or None")
        self.environment = environment
        self._changed()
        return environment

    # -- views -------------------------------------------------------------
    @property
    def instances(self):
        return tuple(self.meshes)

    @property
    def visible_meshes(self):
        return tuple(mesh for mesh in self.meshes if mesh.visible)

    @property
    def render_meshes(self):
        return self.visible_meshes

    visible_volumes = ()

    @property
    def triangle_count(self):
        return sum(len(mesh.indices) for mesh in self.visible_meshes)

    # -- packers (reference layouts) ------------------------------------------
    def triangles(self):
        """World-space triangles, per-triangle base colour (the face material's
        when the mesh has one) and emission."""
        tris, colors, emissions = [], [], []
        for mesh in self.visible_meshes:
            world = mesh.world_vertices[mesh.indices]
            tris.append(world)
            count = len(world)
            faces = mesh.face_materials
            if faces is not None:
                colors.append(faces.base_color)
            else:
                colors.append(np.broadcast_to(np.asarray(mesh.material.base_color, np.float32), (count, 3)))
            emissions.append(np.broadcast_to(np.asarray(mesh.material.emission, np.float32), (count, 3)))
        if not tris:
            empty3 = np.empty((0, 3), np.float32)
            return np.empty((0, 3, 3), np.float32), empty3, empty3.copy()
        return (np.ascontiguousarray(np.concatenate(tris), np.float32),
                np.ascontiguousarray(np.concatenate(colors), np.float32),
                np.ascontiguousarray(np.concatenate(emissions), np.float32))

    def render_triangles(self):
        return self.triangles()[0]

    def triangle_instance_ids(self):
        values = [np.full(len(m.indices), m.id, np.uint32)
                  for m in self.render_meshes if len(m.indices)]
        return np.concatenate(values) if values else np.empty(0, np.uint32)

    triangle_object_ids = triangle_instance_ids

    # -- updates (scene/_core.py:1579-1640, :1809-1850) ------------------------
    def update_mesh_resource(self, resource, *, vertices=None, normals=None):
        """New vertex positions (same count) and normals of shared geometry;
        normals are regenerated when only vertices are given."""
        if resource not in self.mesh_resources:
            raise KeyError("resource does not belong to this scene")
        v = resource.vertices if vertices is None else np.ascontiguousarray(vertices, np.float32)
        if v.shape != resource.vertices.shape:
            raise ValueError("vertices must keep the resource's vertex count")
        if not np.all(np.isfinite(v)):
            raise ValueError("vertices must be finite")
        if normals is None:
            n = generate_normals(v, resource.indices) if vertices is not None else resource.normals
        else:
            n = np.ascontiguousarray(normals, np.float32)
            if n.shape != v.shape:
                raise ValueError("normals must have the same shape as vertices")
            n = n / np.linalg.norm(n, axis=1, keepdims=True)
        resource.vertices = v
        resource.normals = n
        self._changed()
        return resource

    def update_instance(self, instance, *, transform=None, material=None):
        """New placement or material of an instance, keeping its id."""
        instance = self.get_instance(instance)
        if transform is not None:
            instance.transform = transform if isinstance(transform, Transform) else Transform(transform)
        if material is not None:
            if not isinstance(material, Material):
                raise TypeError("material must be a Material")
            self._register_material(material)
            instance.material = material
        self._changed()
        return instance

    def update_instance_transforms(self, instances, transforms):
        """Column update of many instances' transforms."""
        instances = list(instances)
        transforms = list(transforms)
        if len(instances) != len(transforms):
            raise ValueError("instances and transforms must have the same length")
        for inst, t in zip(instances, transforms):
            inst = self.get_instance(inst)
            inst.transform = t if isinstance(t, Transform) else Transform(t)
        self._changed()
        return instances

    def get_instance(self, reference):
        """The instance with this stable id (or the instance itself)."""
        object_id = int(reference.id if hasattr(reference, "id") else reference)
        for mesh in self.meshes:
            if mesh.id == object_id:
                return mesh
        raise KeyError(f"scene has no instance with id {object_id}")

    def object_triangle_range(self, reference):
        """Half-open range of the object's triangles in packed order."""
        object_id = int(reference.id if hasattr(reference, "id") else reference)
        offset = 0
        for mesh in self.render_meshes:
            count = len(mesh.indices)
            if mesh.id == object_id:
                return offset, offset + count
            offset += count
        raise KeyError(f"scene has no visible object with id {object_id}")

    def triangle_material_ids(self):
        values = [np.full(len(m.indices), self.material_id(m.material), np.uint32)
                  for m in self.render_meshes if len(m.indices)]
        return np.concatenate(values) if values else np.empty(0, np.uint32)

    def _emissive_triangles(self):
        """(triangles, emission, area, two_sided) of emissive, non-degenerate triangles."""
        rows = []
        for mesh in self.visible_meshes:
            emission = np.asarray(mesh.material.emission, np.float32)
            if not np.any(emission > 0.0):
                continue
            tri = mesh.world_vertices[mesh.indices]
            area = 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1)
            keep = area > 1e-12
            rows.append((tri[keep], emission, area[keep],
                         float(mesh.material.emission_two_sided)))
        return rows

    def emissive_triangle_data(self):
        """5 x vec4 per emissive triangle: a, b, c, (emission, area), (cdf, pmf, two-sided, 0)."""
        rows = self._emissive_triangles()
        if not rows:
            packed = np.zeros((1, 5, 4), np.float32)
            packed[0, 4, :2] = 1.0
            return packed
        tris = np.concatenate([r[0] for r in rows])
        emission = np.concatenate([np.broadcast_to(r[1], (len(r[0]), 3)) for r in rows])
        area = np.concatenate([r[2] for r in rows])
        sided = np.concatenate([np.full(len(r[0]), r[3], np.float32) for r in rows])
        # Each weight is a float32 product (area * luminance) as in the
        # reference's scalar loop; normalisation and the CDF run in float64.
        weights = np.concatenate([
            (r[2].astype(np.float32) * np.float32(float(np.dot(r[1], _LUMINANCE)))).astype(np.float64)
            for r in rows
        ])
        probabilities = weights / weights.sum()
        cdf = np.cumsum(probabilities)
        cdf[-1] = 1.0
        packed = np.zeros((len(tris), 5, 4), np.float32)
        packed[:, :3, :3] = tris
        packed[:, 3, :3] = emission
        packed[:, 3, 3] = area
        packed[:, 4, 0] = cdf
        packed[:, 4, 1] = probabilities
        packed[:, 4, 2] = sided
        return packed

    @property
    def emissive_light_weight(self):
        """Sum over emissive triangles of area x luminance (float64)."""
        total = 0.0
        for mesh in self.visible_meshes:
            emission = np.asarray(mesh.material.emission, np.float64)
            luminance = float(np.dot(emission, _LUMINANCE))
            if luminance <= 0.0:
                continue
            tri = mesh.world_vertices[mesh.indices]
            areas = 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1)
            total += float(areas[areas > 1e-12].sum()) * luminance
        return total

    @property
    def emissive_triangle_count(self):
        count = 0
        for mesh in self.visible_meshes:
            if not any(c > 0.0 for c in mesh.material.emission):
                continue
            tri = mesh.world_vertices[mesh.indices]
            count += int(np.count_nonzero(np.linalg.norm(
                np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1) > 2e-12))
        return count

    def analytic_light_data(self):
        """4 x vec4 per analytic light: position/type, direction/range, color/intensity, spot/texture."""
        records = []
        for light in self.lights:
            if isinstance(light, PointLight):
                records.append(((*light.position, float(POINT)),
                                (0.0, 0.0, 0.0, light.range or 0.0),
                                (*light.color, light.intensity), (0.0,) * 4))
            elif isinstance(light, DirectionalLight):
                d = np.asarray(light.direction, np.float32)
                d = d / np.linalg.norm(d)
                records.append(((0.0, 0.0, 0.0, float(DIRECTIONAL)), (*d, 0.0),
                                (*light.color, light.intensity), (0.0,) * 4))
            else:
                d = np.asarray(light.direction, np.float32)
                d = d / np.linalg.norm(d)
                records.append(((*light.position, float(SPOT)), (*d, light.range or 0.0),
                                (*light.color, light.intensity),
                                (float(np.cos(light.inner_cone_angle)),
                                 float(np.cos(light.outer_cone_angle)), 0.0, 0.0)))
        if self.environment is not None:
            records.append(((0.0, 0.0, 0.0, float(ENVIRONMENT)), (0.0,) * 4,
                            (*self.environment.color, self.environment.intensity),
                            (-1.0, self.environment.rotation, 0.0, 0.0)))
        if not records:
            records = [((0.0,) * 4,) * 4]
        return np.ascontiguousarray(records, dtype=np.float32).reshape((-1, 4, 4))

    @property
    def analytic_light_count(self):
        return len(self.lights) + int(self.environment is not None)

    def bounds(self):
        tris = self.triangles()[0]
        if not len(tris):
            return np.zeros(3, np.float32), np.zeros(3, np.float32)
        flat = tris.reshape(-1, 3)
        return flat.min(axis=0), flat.max(axis=0)


__all__ = [
    "Texture", "TextureTransform", "MATERIAL_TEXTURE_SLOTS", "generate_tangents",
    "Instance", "Material", "Mesh", "MeshResource", "Scene", "Transform",
    "generate_normals", "material_record",
]
