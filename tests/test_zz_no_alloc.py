"""GPU: olhip_trace_device, the closest-hit query that allocates nothing.

Its hits must equal olhip_trace's (the allocating entry point) bit for bit
on the same rays, for a flat and a two-level scene, on two calls in a row
(the second with fewer rays, on a stream of the caller's, so stale contents
of the shared scratch array or of the status word would show), and its
status word must read zero.  Device buffers come from the HIP runtime
through ctypes; the library call itself allocates nothing.
"""

import ctypes
import sys

import numpy as np
import pytest

from ordinarylight_amd import _native as N
from ordinarylight_amd import workloads
from ordinarylight_amd.renderers.hip import DeviceContext, DeviceScene

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class Hip:
    """The few HIP runtime calls the test needs for device buffers."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        self.rt.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.rt.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        self.rt.hipStreamSynchronize.argtypes = [ctypes.c_void_p]

    def check(self, status, what):
        assert status == 0, f"{what} failed with HIP status {status}"

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        self.check(self.rt.hipMalloc(ctypes.byref(p), nbytes), "hipMalloc")
        return p

    def upload(self, array):
        array = np.ascontiguousarray(array)
        p = self.malloc(array.nbytes)
        self.check(self.rt.hipMemcpy(p, array.ctypes.data, array.nbytes, H2D), "hipMemcpy")
        return p

    def download(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        self.check(self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H), "hipMemcpy")
        return out


@pytest.fixture(scope="module")
def ctx(hip_device):
    c = DeviceContext(hip_device)
    yield c
    c.close()


def rays(scene_box, n, seed):
    lo, hi = scene_box
    rng = np.random.default_rng(seed)
    o = np.empty((n, 4), np.float32)
    d = np.empty((n, 4), np.float32)
    o[:, :3] = rng.uniform(lo - 2, hi + 2, (n, 3))
    o[:, 3] = 0.001
    v = rng.normal(size=(n, 3))
    d[:, :3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    d[:, 3] = 1e30
    return o, d


def allocating_hits(ds, o, d):
    hits = np.empty((len(o), 4), np.float32)
    N.check(N.library().olhip_trace(ds.context.handle, ds.handle, N.ptr(o), N.ptr(d), len(o), N.ptr(hits)))
    return hits


@pytest.mark.parametrize("kind", ["flat", "instanced"])
def test_trace_device_equals_allocating_trace(ctx, kind):
    hip = Hip()
    if kind == "flat":
        scene, _ = workloads.terrain(3, nx=201, nz=101, chunk=50, emissive_quads=4)
        ds = DeviceScene(ctx, scene, 2, builder="sah")
        pos = ds.packed.positions
    else:
        scene, _ = workloads.instanced_spheres(2, meshes=3, subdivisions=2, instances=80, box=12.0,
                                              dielectric=0.2, emitters=20, emitter_size=0.3)
        ds = DeviceScene(ctx, scene, 2, instanced=True, builder="sah")
        pos = np.array([[-8.0, -8.0, -8.0], [8.0, 8.0, 8.0]], np.float32)
    box = (pos.min(0), pos.max(0))
    lib = N.library()
    stream = ctypes.c_void_p()
    hip.check(hip.rt.hipStreamCreate(ctypes.byref(stream)), "hipStreamCreate")
    # Two calls in a row, the second smaller and on the caller's stream.
    for call, (n, seed, strm) in enumerate(((120000, 11, None), (37000, 12, stream))):
        log(f"trace_device {kind} call {call}: {n} rays, stream {'caller' if strm else 'context'}")
        o, d = rays(box, n, seed)
        expect = allocating_hits(ds, o, d)
        d_o, d_d = hip.upload(o), hip.upload(d)
        d_h = hip.malloc(16 * n)
        d_s = hip.malloc(4)
        # Poison the outputs: the call must write every hit and clear the status.
        hip.check(hip.rt.hipMemset(d_h, 0xFF, 16 * n), "hipMemset")
        hip.check(hip.rt.hipMemset(d_s, 0x5A, 4), "hipMemset")
        hip.check(hip.rt.hipDeviceSynchronize(), "hipDeviceSynchronize")
        N.check(lib.olhip_trace_device(ctx.handle, ds.handle, d_o, d_d, n, d_h, d_s, strm))
        if strm is None:
            hip.check(hip.rt.hipDeviceSynchronize(), "hipDeviceSynchronize")
        else:
            hip.check(hip.rt.hipStreamSynchronize(strm), "hipStreamSynchronize")
        got = hip.download(d_h, (n, 4), np.float32)
        status = hip.download(d_s, (1,), np.uint32)[0]
        for p in (d_o, d_d, d_h, d_s):
            hip.check(hip.rt.hipFree(p), "hipFree")
        hit = expect[:, 3].view(np.uint32) != 0xFFFFFFFF
        log(f"  status {status}, hit fraction {hit.mean():.3f}")
        assert status == 0
        assert hit.mean() > 0.05
        assert np.array_equal(got.view(np.uint32), expect.view(np.uint32)), \
            f"call {call}: {np.sum(np.any(got.view(np.uint32) != expect.view(np.uint32), axis=1))} hits differ"
    hip.check(hip.rt.hipStreamDestroy(stream), "hipStreamDestroy")
    ds.close()


def test_trace_device_rejects_null_arguments(ctx):
    scene, _ = workloads.box_room(0)
    ds = DeviceScene(ctx, scene)
    with pytest.raises(ValueError):
        N.check(N.library().olhip_trace_device(ctx.handle, ds.handle, None, None, 4, None, None, None))
    ds.close()
