"""Output-stage oracle -- TEST INFRASTRUCTURE ONLY.

float32 NumPy restatement of the reference's presentation passes:
wavefront_tone_map.comp (ACES approximation :45-48, sRGB encode :50-56,
packUnorm4x8 of (rgb, 1)), rgba_to_nv12.comp (byte values :35-38, luma
:40-43, chroma :45-50, edge clamp :52-55, 4x2 blocks and plane layout
:62-93) and hdr_to_p010.comp (10-bit values :42-45, luma/chroma :46-55,
nearest source pixel :56-59, half-float source image).  Integer rounding is
round-half-even, the choice this build makes for GLSL's round().
Pinned by tests/test_output_oracle.py against closed-form values (black,
white and grey frames, the sRGB breakpoint, ACES limits) and BT.709 studio
ranges.
"""

import numpy as np

F = np.float32


def aces(c):
    c = c.astype(F)
    return np.clip((c * ((F(2.51) * c) + F(0.03))) / ((c * ((F(2.43) * c) + F(0.59))) + F(0.14)), F(0), F(1))


def srgb_encode(c):
    c = c.astype(F)
    lower = c * F(12.92)
    with np.errstate(invalid="ignore"):
        upper = (F(1.055) * np.power(c, F(1.0) / F(2.4))) - F(0.055)
    return np.where(c <= F(0.0031308), lower, upper).astype(F)


def display(hdr, exposure):
    return srgb_encode(aces(np.maximum(hdr[..., :3].astype(F), F(0)) * F(exposure)))


def rgba8(hdr, exposure=1.0):
    """uint8 [H, W, 4] tone-mapped frame, alpha 255."""
    c = display(hdr, exposure)
    out = np.empty(hdr.shape[:2] + (4,), np.uint8)
    out[..., :3] = np.rint(np.clip(c, F(0), F(1)) * F(255)).astype(np.uint8)
    out[..., 3] = 255
    return out


def _luma(c):
    return (c[..., 0] * F(0.2126) + c[..., 1] * F(0.7152)) + c[..., 2] * F(0.0722)


def _cb(c):
    return (c[..., 0] * F(-0.114572) + c[..., 1] * F(-0.385428)) + c[..., 2] * F(0.5)


def _cr(c):
    return (c[..., 0] * F(0.5) + c[..., 1] * F(-0.454153)) + c[..., 2] * F(-0.045847)


def _block_average(c):
    """0.25 * (((top0 + top1) + bottom0) + bottom1) over 2x2 blocks."""
    return F(0.25) * (((c[0::2, 0::2] + c[0::2, 1::2]) + c[1::2, 0::2]) + c[1::2, 1::2])


def nv12(hdr, exposure=1.0):
    """uint8 [H * 3 / 2, W]: luma plane, then interleaved Cb/Cr rows."""
    H, W = hdr.shape[:2]
    c = (rgba8(hdr, exposure)[..., :3].astype(F) / F(255)).astype(F)
    byte = lambda v: np.clip(np.rint(v), 0, 255).astype(np.uint8)
    y = byte(F(16) + F(219) * _luma(c))
    avg = _block_average(c)
    uv = np.empty((H // 2, W), np.uint8)
    uv[:, 0::2] = byte(F(128) + F(224) * _cb(avg))
    uv[:, 1::2] = byte(F(128) + F(224) * _cr(avg))
    return np.concatenate([y, uv])


def p010(hdr, width, height, exposure=1.0):
    """uint16 [height * 3 / 2, width] from the half-float HDR frame."""
    SH, SW = hdr.shape[:2]
    xs = np.clip((np.arange(width) * SW) // width, 0, SW - 1)
    ys = np.clip((np.arange(height) * SH) // height, 0, SH - 1)
    src = hdr[ys][:, xs, :3].astype(np.float16).astype(F)
    c = display(src, exposure)
    ten = lambda v: (np.clip(np.rint(v), 0, 1023).astype(np.uint16) << 6).astype(np.uint16)
    y = ten(F(64) + F(876) * _luma(c))
    avg = _block_average(c)
    uv = np.empty((height // 2, width), np.uint16)
    uv[:, 0::2] = ten(F(512) + F(896) * _cb(avg))
    uv[:, 1::2] = ten(F(512) + F(896) * _cr(avg))
    return np.concatenate([y, uv])
