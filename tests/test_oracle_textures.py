"""CPU: texture packing and the GI oracle's texture sampling, pinned by
closed forms.

* Packing (ordinarylight_amd/textures.py, the layout of the reference's
  Scene.texture_data, scene/_core.py:3005-3043): mip levels are 2 x 2 box
  averages rounded half up (edge repeated for odd sizes), sRGB chains are
  averaged in linear light and re-encoded, and the header points at both
  chains.
* Sampler (wavefront_shade_candidate.glsl:687-827): textbook bilinear and
  nearest filtering on a two-texel texture, repeat / clamp / mirror
  addressing, sRGB decoding, the mip level log2(footprint x size) including
  the binding's uv scale, and the KHR_texture_transform rotation.
* Environment image (:1835-1894): a uniform image returns its radiance up to
  the 8-bit log encoding; a bright texel is found along the direction the
  longitude / colatitude mapping names, and rotation turns it.
* Normal mapping (:889-955): on a plane whose tangent frame is known in
  closed form, a tangent-space normal maps to T*x + B*y + N*z (with the
  handedness from the uv orientation), and a rotated binding turns the
  frame.
* Material application (:915-934): constant textures scale base colour,
  roughness, metallic, emission, transmission and occlusion as stated, and
  an end-to-end render with constant textures equals the untextured render
  of the scaled material.
Each check fails on the planted mistake it pins (OL_MUTANT 13-16).
"""

import math

import numpy as np
import pytest

from ordinarylight_amd import cameras, packing, textures as tx
from ordinarylight_amd.lights import EnvironmentLight
from ordinarylight_amd.scene import Material, Scene, Texture, TextureTransform
from oracle import native as oracle


def srgb_decode(b):
    v = b / 255.0
    return v / 12.92 if v <= 0.04045 else ((v + 0.055) / 1.055) ** 2.4


def plane_scene(material, env=None, uv_scale=1.0):
    """One quad on y = 0 facing +y, uv = ((x + 1) / 2, (z + 1) / 2) * scale."""
    s = Scene()
    v = np.array(((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1)), np.float32)
    i = np.array(((0, 1, 2), (0, 2, 3)), np.uint32)
    uv = np.column_stack(((v[:, 0] + 1) / 2, (v[:, 2] + 1) / 2)).astype(np.float32) * np.float32(uv_scale)
    s.add_mesh(v, i, material, texcoords=uv, texcoords1=uv)
    if env is not None:
        s.set_environment(env)
    return s


def texture(pixels, **kw):
    px = np.asarray(pixels, np.uint8)
    if px.shape[-1] == 3:
        px = np.concatenate((px, np.full(px.shape[:-1] + (1,), 255, np.uint8)), axis=-1)
    return Texture(px, **kw)


def oracle_for(scene, mutant=0, precision="f32"):
    return oracle.OracleScene(packing.pack_scene(scene), precision=precision, mutant=mutant)


# ---- packing --------------------------------------------------------------------------------------
def test_mip_chain_closed_forms():
    one = np.array([[[10, 20, 30, 40]]], np.uint8)
    assert [lv.tolist() for lv in tx.mip_chain(one, False)] == [one.tolist()]
    # 2 x 2 -> 1 x 1: the mean 127.5 rounds half up to 128.
    sq = np.array([[[0] * 4, [255] * 4], [[255] * 4, [0] * 4]], np.uint8)
    levels = tx.mip_chain(sq, False)
    assert len(levels) == 2 and levels[1].tolist() == [[[128] * 4]]
    # sRGB: black and white average to linear 0.5, re-encoded 1.055*0.5^(1/2.4)-0.055.
    bw = np.array([[[0, 0, 0, 255], [255, 255, 255, 255]]], np.uint8)
    enc = 1.055 * 0.5 ** (1 / 2.4) - 0.055
    lv = tx.mip_chain(bw, True)[1][0, 0]
    assert lv[:3].tolist() == [int(enc * 255 + 0.5)] * 3 and lv[3] == 255
    # Odd width: the last column repeats: [a, b, c] -> [(a+b)/2, c] -> [((a+b)/2 + c)/2].
    row = np.array([[[0] * 4, [100] * 4, [200] * 4]], np.uint8)
    levels = tx.mip_chain(row, False)
    assert [lv.shape[:2] for lv in levels] == [(1, 3), (1, 2), (1, 1)]
    assert levels[1][0, :, 0].tolist() == [50, 200]
    assert levels[2][0, 0, 0] == 125


def test_texture_words_layout():
    a = texture(np.arange(3 * 5 * 3).reshape(3, 5, 3) % 256, wrap_s="mirror", wrap_t="clamp", linear_filter=False)
    b = texture(np.full((1, 1, 3), 7))
    w = tx.texture_words([a, b])
    assert w[0] == 2
    sizes_a = [5 * 3, 3 * 2, 2 * 1, 1]
    meta_a, meta_b = w[1:9], w[9:17]
    assert meta_a[2:6].tolist() == [5, 3, 2 | (1 << 2) | (0 << 4), 4]
    assert meta_a[0] == 17 and meta_a[1] == 17 + sum(sizes_a)
    assert meta_b[0] == 17 + 2 * sum(sizes_a) and meta_b[1] == meta_b[0] + 1
    assert len(w) == meta_b[1] + 1
    # Texels are RGBA bytes, R in the low byte; the linear chain keeps the bytes.
    first = w[meta_a[1]]
    assert [(first >> s) & 255 for s in (0, 8, 16, 24)] == a.pixels[0, 0].tolist()


# ---- sampler ----------------------------------------------------------------------------------------
BW = [[[0, 0, 0], [255, 255, 255]]]


@pytest.mark.parametrize("mutant", [0, 13])
def test_bilinear_nearest_and_wrapping(mutant):
    cases = [
        (dict(), [(0.5, 0.5), (0.25, 1.0 / 3.0), (0.75, 0.9), (0.0, 0.5)], [0.5, 0.0, 1.0, 0.5]),
        (dict(wrap_s="clamp"), [(0.0, 0.5), (1.3, 0.5)], [0.0, 1.0]),
        (dict(wrap_s="mirror"), [(1.25, 0.5), (-0.25, 0.5)], [1.0, 0.0]),
        (dict(linear_filter=False), [(0.49, 0.5), (0.51, 0.5), (-0.2, 0.5)], [0.0, 1.0, 1.0]),
    ]
    ok = True
    for sampler, uvs, want in cases:
        o = oracle_for(plane_scene(Material(base_color_texture=texture(BW, **sampler))), mutant)
        got = [float(o.texture_sample(0, uv)[0]) for uv in uvs]
        o.close()
        ok = ok and np.allclose(got, want, atol=1e-6)
        if mutant == 0:
            assert np.allclose(got, want, atol=1e-6), (sampler, got, want)
    if mutant:
        assert not ok, "the planted mistake went unnoticed"


@pytest.mark.parametrize("mutant", [0, 14])
def test_srgb_decoding(mutant):
    o = oracle_for(plane_scene(Material(base_color_texture=texture([[[188, 60, 10]]]))), mutant)
    got = o.texture_sample(0, (0.5, 0.5), srgb=True)[:3]
    raw = o.texture_sample(0, (0.5, 0.5), srgb=False)[:3]
    o.close()
    want = np.array([srgb_decode(b) for b in (188, 60, 10)])
    assert np.allclose(raw, np.array([188, 60, 10]) / 255.0, atol=1e-7)
    if mutant == 0:
        assert np.allclose(got, want, rtol=1e-5, atol=1e-7)
    else:
        assert not np.allclose(got, want, rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("mutant", [0, 15])
def test_mip_level_from_footprint(mutant):
    px = [[[0, 0, 0], [0, 0, 0]], [[0, 0, 0], [255, 255, 255]]]   # level 1 = 63.75 -> 64
    tf = TextureTransform(scale=(4.0, 4.0))
    o = oracle_for(plane_scene(Material(base_color_texture=texture(px), base_color_transform=tf)), mutant)
    # uv (1/16, 1/16) is texel (0, 0)'s centre after the x4 scale.
    at = (1.0 / 16.0, 1.0 / 16.0)
    level0 = float(o.texture_sample(0, at, footprint0=0.0)[0])
    level1 = float(o.texture_sample(0, at, footprint0=0.25)[0])     # 0.25 x scale 4 x size 2 = 2 -> lod 1
    half = float(o.texture_sample(0, at, footprint0=0.25 / math.sqrt(2.0))[0])  # lod 0.5
    o.close()
    checks = [abs(level0) < 1e-7, abs(level1 - 64 / 255) < 1e-6, abs(half - 0.5 * (64 / 255)) < 1e-6]
    assert all(checks) if mutant == 0 else not all(checks)


def test_binding_rotation_and_offset():
    px = [[[255, 0, 0], [0, 255, 0]], [[0, 0, 255], [255, 255, 0]]]
    rot = TextureTransform(rotation=math.pi / 2)
    off = TextureTransform(offset=(0.5, 0.0), scale=(1.0, 1.0))
    tex = texture(px, linear_filter=False)
    m = Material(base_color_texture=tex, base_color_transform=rot, emissive_texture=tex, emissive_transform=off,
                 metallic_roughness_texture=tex)
    o = oracle_for(plane_scene(m))
    # KHR_texture_transform: uv' = offset + R(theta) * (scale * uv); a quarter
    # turn sends (u, v) to (-v, u), so (0.25, 0.75) reads (-0.75, 0.25) = texel (0, 0).
    # Bindings in slot order: base colour 0, metallic-roughness 1 (identity), emissive 2.
    plain = o.texture_sample(1, (0.25, 0.25))
    turned = o.texture_sample(0, (0.25, 0.75))
    shifted = o.texture_sample(2, (0.25, 0.25))               # + (0.5, 0): texel (1, 0)
    o.close()
    assert np.allclose(plain[:3], (1, 0, 0)) and np.allclose(turned[:3], (1, 0, 0))
    assert np.allclose(shifted[:3], (0, 1, 0))


# ---- environment ------------------------------------------------------------------------------------
def test_environment_image_uniform_and_direction():
    c = np.array((2.0, 1.0, 0.5), np.float32)
    env = EnvironmentLight(image=np.broadcast_to(c, (4, 8, 3)).copy(), color=(0.5, 1.0, 1.0), intensity=2.0)
    o = oracle_for(plane_scene(Material(), env=env))
    for d in ((0, 1, 0), (1, 0, 0), (0.3, -0.5, 0.81)):
        d = np.asarray(d, np.float32) / np.linalg.norm(d)
        got = o.environment_radiance(d)
        assert np.allclose(got, c * np.array((0.5, 1.0, 1.0)) * 2.0, rtol=0.01), got
    o.close()
    # One bright texel block (columns 4, rows 1-2 of 8 x 4): found along
    # longitude (u - 0.5) * 2 pi from +x towards +z, colatitude pi / 2, and
    # turned by the rotation.
    img = np.full((4, 8, 3), 0.01, np.float32)
    img[1:3, 4] = 100.0
    for rotation in (0.0, 0.9):
        o = oracle_for(plane_scene(Material(), env=EnvironmentLight(image=img, rotation=rotation)))
        lon = (4.5 / 8 - 0.5) * 2 * math.pi - rotation
        bright = o.environment_radiance(np.array((math.cos(lon), 0.0, math.sin(lon)), np.float32))
        dark = o.environment_radiance(np.array((-math.cos(lon), 0.0, -math.sin(lon)), np.float32))
        o.close()
        assert bright[0] > 90.0 and dark[0] < 0.05, (rotation, bright, dark)


# ---- normal mapping ---------------------------------------------------------------------------------
def encode_normal(n):
    n = np.asarray(n, np.float64)
    return np.clip(np.floor((n * 0.5 + 0.5) * 255 + 0.5), 0, 255).astype(np.uint8)


def expected_normal(byte_rgb, frame):
    t, b, n = (np.asarray(v, np.float64) for v in frame)
    ts = np.asarray(byte_rgb, np.float64) / 255.0 * 2.0 - 1.0
    ts /= np.linalg.norm(ts)
    w = t * ts[0] + b * ts[1] + n * ts[2]
    return w / np.linalg.norm(w)


@pytest.mark.parametrize("mutant", [0, 16])
def test_normal_map_tangent_frame(mutant):
    # Plane y = 0, normal +y, u along +x, v along +z: tangent +x; the
    # uv-gradient bitangent is +z while n x t = -z, so the handedness is -1
    # and B = (n x t) * -1 = +z.
    frame = ((1, 0, 0), (0, 0, 1), (0, 1, 0))
    ok = True
    for ts in ((0.6, 0.0, 0.8), (0.0, 0.6, 0.8), (-0.3, 0.5, 0.81)):
        px = encode_normal(np.array(ts) / np.linalg.norm(ts))[None, None]
        m = Material(normal_texture=texture(px))
        o = oracle_for(plane_scene(m), mutant)
        got = o.textured_surface(0, 0.25, 0.25)["normal"]
        o.close()
        want = expected_normal(px[0, 0], frame)
        ok = ok and np.allclose(got, want, atol=2e-6)
        if mutant == 0:
            assert np.allclose(got, want, atol=2e-6), (ts, got, want)
    if mutant:
        assert not ok
    # A quarter-turned binding turns the frame: T' = -B, B' = T.
    px = encode_normal((0.6, 0.0, 0.8))[None, None]
    o = oracle_for(plane_scene(Material(normal_texture=texture(px),
                                        normal_transform=TextureTransform(rotation=math.pi / 2))))
    got = o.textured_surface(1, 0.3, 0.3)["normal"]
    o.close()
    assert np.allclose(got, expected_normal(px[0, 0], ((0, 0, -1), (1, 0, 0), (0, 1, 0))), atol=2e-6)


# ---- material application ---------------------------------------------------------------------------
def test_material_textures_scale_the_material():
    base = texture([[[188, 120, 40]]])
    mr = texture([[[0, 128, 64]]])
    occ = texture([[[100, 0, 0]]])
    emit = texture([[[200, 100, 50]]])
    m = Material(base_color=(0.5, 0.6, 0.7), roughness=0.8, metallic=0.9, emission=(2.0, 3.0, 4.0),
                 base_color_texture=base, metallic_roughness_texture=mr, occlusion_texture=occ,
                 occlusion_strength=0.7, emissive_texture=emit)
    o = oracle_for(plane_scene(m))
    got = o.textured_surface(0, 0.2, 0.3)
    o.close()
    dec = lambda px: np.array([srgb_decode(b) for b in px])  # noqa: E731
    assert np.allclose(got["base"], np.array((0.5, 0.6, 0.7)) * dec((188, 120, 40)), rtol=2e-6)
    assert np.isclose(got["roughness"], 0.8 * 128 / 255, rtol=2e-6)
    assert np.isclose(got["metallic"], 0.9 * 64 / 255, rtol=2e-6)
    assert np.isclose(got["occlusion"], 0.3 + 0.7 * 100 / 255, rtol=2e-6)
    assert np.allclose(got["emission"], np.array((2.0, 3.0, 4.0)) * dec((200, 100, 50)), rtol=2e-6)
    # Transmissive after its texture: metallic-roughness and occlusion are not applied.
    g = Material(transmission=1.0, roughness=0.8, transmission_texture=texture([[[200, 0, 0]]]),
                 metallic_roughness_texture=mr)
    o = oracle_for(plane_scene(g))
    got = o.textured_surface(0, 0.2, 0.3)
    o.close()
    assert np.isclose(got["transmission"], 200 / 255, rtol=2e-6) and np.isclose(got["roughness"], 0.8)


def test_constant_textures_render_as_scaled_material():
    """End to end: a scene whose textures are constant renders like the
    untextured scene with the material scaled by the decoded texels.  (An
    emissive texture is left out: as in the reference, next-event estimation
    samples emitters by their untextured emission, so it changes the image
    differently from scaled emission.)"""
    from ordinarylight_amd import workloads
    scene, cam = workloads.box_room(3)
    base_b, mr_b = (170, 90, 200), (0, 140, 30)
    textured, plain = Scene(), Scene()
    f32 = np.float32
    for mesh in scene.meshes:
        m = mesh.material
        args = dict(base_color=m.base_color, emission=m.emission, roughness=0.7, metallic=0.4)
        textured.add_mesh(mesh.world_vertices, mesh.indices, Material(
            **args, base_color_texture=texture([[base_b]]), metallic_roughness_texture=texture([[mr_b]])),
            texcoords=np.zeros((len(mesh.vertices), 2), np.float32))
        dec = np.array([srgb_decode(b) for b in base_b], np.float32)
        plain.add_mesh(mesh.world_vertices, mesh.indices, Material(
            base_color=tuple(float(v) for v in np.float32(m.base_color) * dec), emission=m.emission,
            roughness=float(f32(0.7) * (f32(140) / f32(255))), metallic=float(f32(0.4) * (f32(30) / f32(255)))))
    images = []
    for s in (textured, plain):
        o = oracle.OracleScene(packing.pack_scene(s))
        images.append(o.render(cameras.camera_basis(cam), cameras.camera_projection(cam), 40, 30, 4, 0,
                               max_bounces=4))
        o.close()
    a, b = images[0][..., :3], images[1][..., :3]
    close = np.all(np.abs(a - b) <= 1e-6 + 1e-4 * np.maximum(np.abs(a), np.abs(b)), axis=-1)
    assert close.mean() >= 0.999, close.mean()
    assert abs(a.mean() - b.mean()) <= 1e-4 * b.mean()
