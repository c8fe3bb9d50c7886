"""tr_scene_trace_rays_noalloc (the allocation-free ray-set trace) against
tr_scene_trace_rays on the same rays: identical hits, bit for bit, on two
calls in a row, the second with fewer rays, through the C ABI with device
buffers the test allocates through the HIP runtime the library itself
links (libamdhip64.so.7, already loaded by the library)."""
import ctypes

import numpy as np
import pytest

from tempestrenderer_amd import workloads
from tempestrenderer_amd.scene import HIT_DTYPE, RayTracerScene

from conftest import gpu_material

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / hipMemset / hipFree of the loaded runtime."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipDeviceSynchronize.argtypes = []

    def malloc(self, n):
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), n) == 0
        return p.value

    def h2d(self, dst, arr):
        assert self.rt.hipMemcpy(dst, arr.ctypes.data, arr.nbytes, 1) == 0

    def d2h(self, arr, src):
        assert self.rt.hipMemcpy(arr.ctypes.data, src, arr.nbytes, 2) == 0


@pytest.mark.parametrize("scene", ["box", "heightfield256"])
def test_trace_rays_noalloc_matches_trace_rays(scene):
    print("case: noalloc trace", scene, flush=True)
    desc = workloads.box(64, 64) if scene == "box" else workloads.heightfield(160, 90, grid=256)
    g = RayTracerScene(desc.width, desc.height, desc.vpi)
    desc.populate(g, gpu_material)
    g.commitScene()
    rays = np.concatenate([workloads.primary_rays(desc), workloads.random_rays(desc, 1 << 16)])
    hip = _Hip()
    n0 = rays.shape[0]
    d_rays = hip.malloc(rays.nbytes)
    d_hits = hip.malloc(n0 * HIT_DTYPE.itemsize)
    try:
        for n in (n0, 1000):
            r = np.ascontiguousarray(rays[:n])
            want = g.traceRays(r)
            hip.h2d(d_rays, r)
            assert hip.rt.hipMemset(d_hits, 0xAB, n0 * HIT_DTYPE.itemsize) == 0
            assert hip.rt.hipDeviceSynchronize() == 0
            g.traceRaysNoAlloc(d_rays, d_hits, n)
            got = np.zeros(n, dtype=HIT_DTYPE)
            hip.d2h(got, d_hits)
            for f in ("geom", "prim"):
                assert np.array_equal(got[f], want[f]), (n, f)
            for f in ("t", "u", "v"):
                assert np.array_equal(got[f].view(np.uint32), want[f].view(np.uint32)), (n, f)
            # past the rays of this call the buffer is untouched
            if n < n0:
                raw = np.zeros(n0 * HIT_DTYPE.itemsize, dtype=np.uint8)
                hip.d2h(raw, d_hits)
                assert np.all(raw[n * HIT_DTYPE.itemsize:] == 0xAB)
    finally:
        hip.rt.hipFree(d_rays)
        hip.rt.hipFree(d_hits)
    g.close()


def test_trace_rays_noalloc_checks_arguments():
    desc = workloads.box(16, 16)
    g = RayTracerScene(desc.width, desc.height, desc.vpi)
    desc.populate(g, gpu_material)
    with pytest.raises(Exception):
        g.traceRaysNoAlloc(0, 0, 16)      # null buffers
    g.commitScene()
    g.traceRaysNoAlloc(0, 0, 0)          # no rays: nothing to do
    g.close()
