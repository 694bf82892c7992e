"""CPU: pins of the output-stage oracle (tone map, NV12, P010).

Closed forms, not retyped formulas: a black frame is (0, 0, 0, 255) and
studio-range black (16/128/128, P010 64/512/512 << 6); ACES saturates to 1
so a very bright frame is white and studio-range white (235/128/128, P010
940/512/512 << 6); sRGB is continuous at its breakpoint and encodes 0.5 to
0.7354; the 2x2 chroma of a frame whose columns alternate is the average
colour; P010 scaling picks nearest source pixels.
"""

import numpy as np

from oracle import output_numpy as out


def frame(rgb, h=4, w=8):
    f = np.zeros((h, w, 4), np.float32)
    f[..., :3] = rgb
    f[..., 3] = 1
    return f


def test_black_and_white_frames():
    b = frame((0.0, 0.0, 0.0))
    assert np.all(out.rgba8(b) == [0, 0, 0, 255])
    nv = out.nv12(b)
    assert nv.shape == (6, 8) and np.all(nv[:4] == 16) and np.all(nv[4:] == 128)
    p = out.p010(b, 8, 4)
    assert np.all(p[:4] == 64 << 6) and np.all(p[4:] == 512 << 6)
    w = frame((1e4, 1e4, 1e4))
    assert np.all(out.rgba8(w) == 255)
    nvw = out.nv12(w)
    assert np.all(nvw[:4] == 235) and np.all(nvw[4:] == 128)
    pw = out.p010(w, 8, 4)
    assert np.all(pw[:4] == 940 << 6) and np.all(pw[4:] == 512 << 6)


def test_aces_and_srgb_closed_forms():
    x = np.array([0.0, 1.0, 100.0], np.float32)
    a = out.aces(x)
    assert a[0] == 0.0 and abs(a[1] - 2.54 / 3.16) < 1e-6 and a[2] == 1.0
    # ACES is monotone on [0, inf).
    grid = np.linspace(0, 20, 2001, dtype=np.float32)
    assert np.all(np.diff(out.aces(grid)) >= 0)
    # sRGB: linear segment, continuity at the breakpoint, 0.5 -> 0.7354.
    assert out.srgb_encode(np.float32([0.001]))[0] == np.float32(0.001) * np.float32(12.92)
    lo = out.srgb_encode(np.float32([0.0031308]))[0]
    hi = out.srgb_encode(np.nextafter(np.float32(0.0031308), np.float32(1)))
    assert abs(float(hi) - float(lo)) < 1e-4
    assert abs(out.srgb_encode(np.float32([0.5]))[0] - 0.735357) < 1e-5


def test_chroma_is_block_average_and_exposure_scales():
    f = frame((0.0, 0.0, 0.0))
    f[:, 0::2, :3] = (0.8, 0.1, 0.1)
    f[:, 1::2, :3] = (0.1, 0.1, 0.8)
    nv = out.nv12(f)
    c = out.rgba8(f)[..., :3].astype(np.float32) / 255
    avg = c.reshape(2, 2, 4, 2, 3).mean(axis=(1, 3))
    cb = np.rint(128 + 224 * (avg @ np.float32([-0.114572, -0.385428, 0.5])))
    assert np.allclose(nv[4:, 0::2], cb, atol=1)
    # Exposure multiplies radiance before the curve.
    assert np.array_equal(out.rgba8(frame((0.25, 0.5, 1.0)), 2.0), out.rgba8(frame((0.5, 1.0, 2.0))))


def test_p010_nearest_scaling():
    f = np.zeros((2, 4, 4), np.float32)
    f[0, :, :3] = 0.0
    f[1, :, :3] = 1e4
    p = out.p010(f, 8, 4)  # 2x upscale: rows 0-1 from source row 0, rows 2-3 from row 1
    assert np.all(p[0:2] == 64 << 6) and np.all(p[2:4] == 940 << 6)
