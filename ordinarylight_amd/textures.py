"""Texture packing for the GI renderer.

The device reads textures from one word array in the layout of the
reference's ``Scene.texture_data`` (``scene/_core.py:3005-3043``), kept alike
because the sampling code indexes it the same way (see DESIGN): word 0 holds
the texture count; texture k has eight metadata words at ``1 + 8k``
(sRGB-chain offset, linear-chain offset, width, height, flags, mip levels,
0, 0; flags = wrap_s | wrap_t << 2 | linear_filter << 4 with repeat 0,
clamp 1, mirror 2); then each texture's two RGBA8 mip chains, level 0 first,
one texel per word (R in the low byte).  The sRGB chain is filtered in
linear light and re-encoded, so a fetch decodes it back to linear.

Bindings are ``[B, 2, 4]`` float records (texture index, cos, sin, texcoord
set), (offset xy, scale xy) as ``texture_binding_data`` (``:2984-3003``).

An environment image is stored as the reference does
(``_pack_environment_texture``, ``:2429-2439``): RGBA8 of
``log2(1 + radiance) / log_range``, wrap repeat/clamp, bilinear; shading
decodes ``exp2(encoded * log_range) - 1``.
"""

from __future__ import annotations

import math

import numpy as np

from .scene import Texture

WRAP_CODES = {"repeat": 0, "clamp": 1, "mirror": 2}


def srgb_decode(x):
    """sRGB transfer function to linear (IEC 61966-2-1), elementwise."""
    x = np.asarray(x, np.float32)
    return np.where(x <= 0.04045, x / np.float32(12.92), ((x + np.float32(0.055)) / np.float32(1.055)) ** 2.4)


def srgb_encode(x):
    """Linear to sRGB, elementwise (negative values clamp to zero)."""
    x = np.maximum(np.asarray(x, np.float32), 0.0)
    return np.where(x <= 0.0031308, x * np.float32(12.92), np.float32(1.055) * x ** (1.0 / 2.4) - np.float32(0.055))


def mip_chain(pixels, srgb):
    """RGBA8 levels down to 1 x 1: each level averages 2 x 2 texels of the
    previous one (the last row / column repeated for odd sizes), in linear
    light when ``srgb`` (colour channels decoded, averaged, re-encoded;
    alpha always linear), bytes rounded half up."""
    level = np.asarray(pixels, np.float32) / np.float32(255.0)
    if srgb:
        level = level.copy()
        level[..., :3] = srgb_decode(level[..., :3])
    levels = []
    while True:
        out = level.copy()
        if srgb:
            out[..., :3] = srgb_encode(out[..., :3])
        levels.append(np.clip(out * np.float32(255.0) + np.float32(0.5), 0, 255).astype(np.uint8))
        h, w = level.shape[:2]
        if h == 1 and w == 1:
            return levels
        padded = np.pad(level, ((0, h & 1), (0, w & 1), (0, 0)), mode="edge")
        nh, nw = (h + 1) // 2, (w + 1) // 2
        level = padded.reshape(nh, 2, nw, 2, 4).mean(axis=(1, 3), dtype=np.float32)


def texture_words(textures):
    """The device word array for ``textures`` (in binding-index order)."""
    header = np.zeros(1 + 8 * len(textures), np.uint32)
    header[0] = len(textures)
    chunks = [header]
    offset = len(header)
    for k, tex in enumerate(textures):
        h, w = tex.pixels.shape[:2]
        flags = WRAP_CODES[tex.wrap_s] | (WRAP_CODES[tex.wrap_t] << 2) | (int(bool(tex.linear_filter)) << 4)
        starts = []
        levels = 0
        for srgb in (True, False):
            chain = mip_chain(tex.pixels, srgb)
            levels = len(chain)
            words = np.concatenate([np.ascontiguousarray(lv).reshape(-1, 4).view(np.uint32).reshape(-1)
                                    for lv in chain])
            starts.append(offset)
            chunks.append(words)
            offset += len(words)
        header[1 + 8 * k:1 + 8 * k + 8] = (starts[0], starts[1], w, h, flags, levels, 0, 0)
    return np.ascontiguousarray(np.concatenate(chunks), np.uint32)


def binding_records(bindings, textures):
    """``[B, 2, 4]`` float32 records for (texture, transform) bindings."""
    records = np.zeros((max(len(bindings), 1), 2, 4), np.float32)
    if not bindings:
        records[0] = ((-1.0, 1.0, 0.0, 0.0), (0.0, 0.0, 1.0, 1.0))
        return records
    for b, (tex, tf) in enumerate(bindings):
        index = next(i for i, t in enumerate(textures) if t is tex)
        records[b, 0] = (float(index), math.cos(tf.rotation), math.sin(tf.rotation), float(tf.texcoord_set))
        records[b, 1] = (*tf.offset, *tf.scale)
    return records


def environment_texture(environment):
    """(Texture, log_range) of an image environment, or (None, 0.0)."""
    image = getattr(environment, "image", None) if environment is not None else None
    if image is None:
        return None, 0.0
    image = np.asarray(image, np.float32)[..., :3]
    log_range = max(float(np.log2(1.0 + np.max(image))), 1e-6)
    encoded = np.zeros((*image.shape[:2], 4), np.uint8)
    encoded[..., :3] = np.clip(np.log2(1.0 + image) / log_range * 255.0 + 0.5, 0, 255).astype(np.uint8)
    encoded[..., 3] = 255
    return Texture(encoded, wrap_s="repeat", wrap_t="clamp"), log_range
