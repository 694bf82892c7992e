"""GPU: display outputs (RGBA8, NV12, P010) against the NumPy oracle.

The device evaluates powf and the ACES rational in float32 like the oracle,
but its pow may differ from NumPy's by an ulp, which can move a value across
an 8- or 10-bit rounding boundary: outputs must be exact on >= 99.9 % of
bytes and never differ by more than one code.
"""

import numpy as np
import pytest

from ordinarylight_amd import workloads
from ordinarylight_amd.renderers.hip import HipPathTracer
from oracle import output_numpy as ref

pytestmark = pytest.mark.gpu


def close_codes(got, want, shift=0):
    g = got.astype(np.int64) >> shift
    w = want.astype(np.int64) >> shift
    d = np.abs(g - w)
    return d.max() <= 1 and (d == 0).mean() >= 0.999


@pytest.mark.parametrize("exposure", [1.0, 3.5])
def test_outputs_match_oracle(hip_device, exposure):
    scene, cam = workloads.box_room(1)
    r = HipPathTracer(max_bounces=4, samples_per_pixel=2, device_index=hip_device, wavefront_exposure=exposure)
    W, H = 96, 64
    hdr = r.render_frame(scene, cam, W, H)
    rgba = r.render_output(scene, cam, W, H, pixel_format="rgba8")
    assert rgba.shape == (H, W, 4) and np.all(rgba[..., 3] == 255)
    assert close_codes(rgba, ref.rgba8(hdr, exposure))
    nv = r.render_output(scene, cam, W, H, pixel_format="nv12")
    assert nv.shape == (H * 3 // 2, W) and close_codes(nv, ref.nv12(hdr, exposure))
    p = r.render_output(scene, cam, W, H, pixel_format="p010")
    assert p.shape == (H * 3 // 2, W) and np.all((p & 63) == 0)
    assert close_codes(p, ref.p010(hdr, W, H, exposure), shift=6)
    # Nearest-neighbour P010 scaling (dynamic resolution: 96x64 shown at 192x96).
    ps = r.render_output(scene, cam, W, H, pixel_format="p010", output_size=(192, 96))
    assert ps.shape == (144, 192) and close_codes(ps, ref.p010(hdr, 192, 96, exposure), shift=6)
    with pytest.raises(ValueError):
        r.render_output(scene, cam, 90, 64, pixel_format="nv12")
    r.close()


def _hip_runtime():
    """The HIP runtime the library itself uses (already loaded in this process)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def test_output_into_device_memory(hip_device):
    import ctypes
    scene, cam = workloads.box_room(2)
    r = HipPathTracer(max_bounces=3, device_index=hip_device)
    W, H = 64, 48
    host = r.render_output(scene, cam, W, H, pixel_format="nv12")
    hip = _hip_runtime()
    size = H * 3 // 2 * W
    ptr = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(ptr), size) == 0
    try:
        assert r.render_output(scene, cam, W, H, pixel_format="nv12", device_out=ptr.value) is None
        back = np.zeros((H * 3 // 2, W), np.uint8)
        assert hip.hipMemcpy(back.ctypes.data_as(ctypes.c_void_p), ptr, size, 2) == 0  # device to host
        assert np.array_equal(back, host)
    finally:
        hip.hipFree(ptr)
    r.close()
