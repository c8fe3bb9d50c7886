"""The reference's Cornell-box scenes through the OBJ loader and the drop-in
scene API, GPU against the CPU oracle (the Blinn-Phong materials of
LoadObjFileStaticRTGeometry, illumination-models.hh:618-698).

Contract as for the other path parity tests: per-path radiance within
1e-4 relative with identical hit-sequence hashes.
"""
import os

import numpy as np
import pytest

import oracle
from tempestrenderer_amd import workloads
from tempestrenderer_amd.scene import LoadObjFileStaticRTGeometry, RayTracerScene

pytestmark = pytest.mark.gpu

CORNELL = os.path.join(os.path.dirname(__file__), "golden", "cornell")


def _setup(target, blob, spp, depth):
    target.setSamplesCamera(spp)
    target.setSamplesGlobalIllumination(1)
    target.setSamplesLocalAreaLight(1)
    target.setMaxRayDepth(depth)
    target.setRussianRoulette(1.0)
    target.setBackgroundSpectrum((0.0, 0.0, 0.0))
    target.setSeed(1)
    target.setEstimator(0)
    target.addTriangleMesh(workloads.IDENTITY, blob.Submeshes, blob.IndexData, blob.VertexData)
    target.commitScene()


@pytest.mark.parametrize("name", ["Original", "Glossy"])
def test_cornell_paths(name):
    print("case: cornell", name, flush=True)
    blob = LoadObjFileStaticRTGeometry(os.path.join(CORNELL, "CornellBox-%s.obj" % name))
    w, h = 64, 64
    vpi = workloads.perspective_view_proj_inv((0.0, 1.0, 5.5), (0.0, 1.0, 0.0), (0.0, 1.0, 0.0), 25.0, w / h)
    g = RayTracerScene(w, h, vpi)
    _setup(g, blob, 4, 4)
    o = oracle.OracleScene(w, h, vpi)
    _setup(o, blob, 4, 4)
    px = np.arange(w * h, dtype=np.uint32)
    g_rgb, g_hash = g.renderPaths(px, 4)
    o_rgb, o_hash = o.renderPaths(px)
    assert np.array_equal(g_hash, o_hash), "hit sequences differ on %d paths" % np.sum(g_hash != o_hash)
    tol = 1e-4 * np.maximum(np.maximum(np.abs(g_rgb), np.abs(o_rgb)), 1e-6)
    bad = np.abs(g_rgb - o_rgb) > tol
    assert not bad.any(), "radiance differs on %d of %d values" % (bad.sum(), bad.size)
    assert g_rgb.mean() > 0.0  # the ceiling light is seen
