"""User shapes, rayQuery, render modes and the transparent background on
the GPU, against the CPU oracle through the C ABI.

- closest hits of fixed rays on a scene of every user shape
  (ray-tracing.cc:1315-1446) and on the reference's own virtual
  gonioreflectometer scene (three cylinders,
  samples/virtual-gonioreflectometer.cc:141-143): ids, t, u, v bit-exact
  against the oracle's brute-force loop;
- per-path radiance and hit sequences on both scenes (both estimators, the
  texture-aware and texture-free shading instances), and the
  gonioreflectometer's 81 light directions (its request loop, :159-182)
  as one batch of render requests;
- rayQuery (ray-tracing.cc:1932-1979): ids, frames and TexCoord;
- RenderMode debug outputs (ray-tracing.cc:2078-2089) path by path;
- TransparentBackground (ray-tracing.cc:2364-2365): the backbuffer equals
  the oracle's resolve of the GPU's sums with the oracle's camera-hit mask,
  untouched (0) pixels where no camera ray hit.
"""
import copy

import numpy as np
import pytest

import oracle
from tempestrenderer_amd import workloads
from tempestrenderer_amd._abi import TempestError
from tempestrenderer_amd.scene import RayTracerScene

from conftest import gpu_material

pytestmark = pytest.mark.gpu


def _pair(desc):
    g = RayTracerScene(desc.width, desc.height, desc.vpi)
    desc.populate(g, gpu_material)
    g.commitScene()
    o = oracle.OracleScene(desc.width, desc.height, desc.vpi)
    desc.populate(o, oracle.material)
    o.commitScene()
    return g, o


def _check_paths(g_rgb, o_rgb, g_hash, o_hash):
    assert np.array_equal(g_hash, o_hash), "hit sequences differ on %d paths" % np.sum(g_hash != o_hash)
    tol = 1e-4 * np.maximum(np.maximum(np.abs(g_rgb), np.abs(o_rgb)), 1e-6)
    bad = np.abs(g_rgb - o_rgb) > tol
    assert not bad.any(), "radiance differs on %d of %d values" % (bad.sum(), bad.size)


def _scene(name):
    if name == "shapes":
        return workloads.shapes(40, 40, samples_camera=4, max_depth=4)
    return workloads.gonio_cylinders(30, 30, samples_camera=8)


@pytest.mark.parametrize("name", ["shapes", "gonio_cylinders"])
def test_shape_ids_match_bruteforce(name):
    print("case: shape ids", name, flush=True)
    desc = _scene(name)
    g, o = _pair(desc)
    rays = np.concatenate([workloads.primary_rays(desc), workloads.random_rays(desc, 1 << 16)])
    hb = o.traceRays(rays, brute_force=True)
    hg = g.traceRays(rays)
    assert np.array_equal(hb["geom"], hg["geom"]) and np.array_equal(hb["prim"], hg["prim"])
    hit = hb["geom"] != 0xFFFFFFFF
    assert np.array_equal(hb["t"][hit].view(np.uint32), hg["t"][hit].view(np.uint32))
    assert np.array_equal(hb["u"][hit], hg["u"][hit]) and np.array_equal(hb["v"][hit], hg["v"][hit])
    # every shape is hit
    slots = set(np.unique(hb["geom"][hit]).tolist())
    assert len(slots) >= (12 if name == "shapes" else 3)


@pytest.mark.parametrize("estimator", [0, 1])
@pytest.mark.parametrize("tangent_map", [True, False])
def test_paths_shapes(estimator, tangent_map):
    print("case: paths shapes estimator=%d tangent_map=%d" % (estimator, tangent_map), flush=True)
    desc = workloads.shapes(40, 40, samples_camera=4, max_depth=4, tangent_map=tangent_map)
    desc.settings.update(estimator=estimator)
    g, o = _pair(desc)
    px = np.arange(desc.pixels, dtype=np.uint32)
    g_rgb, g_hash = g.renderPaths(px, 4)
    o_rgb, o_hash = o.renderPaths(px)
    _check_paths(g_rgb, o_rgb, g_hash, o_hash)


def test_paths_gonio_cylinders():
    print("case: paths gonioreflectometer cylinders 30x30 x 64 spp", flush=True)
    desc = workloads.gonio_cylinders(30, 30, samples_camera=64)
    g, o = _pair(desc)
    px = np.arange(desc.pixels, dtype=np.uint32)
    g_rgb, g_hash = g.renderPaths(px, 64)
    o_rgb, o_hash = o.renderPaths(px)
    _check_paths(g_rgb, o_rgb, g_hash, o_hash)
    assert o_rgb.max() > 0.0


def test_gonio_cylinder_light_loop_as_requests():
    """The 81 light directions of the reference's loop, 90x90 pixels, 4
    camera samples each, in one render-requests call against the oracle
    rendering each request in turn."""
    print("case: gonioreflectometer cylinders, 81 requests", flush=True)
    d = workloads.gonio_cylinders(90, 90, samples_camera=4)
    reqs = [dict(view_proj_inv=d.vpi, width=90, height=90, samples_camera=4,
                 light_direction=workloads.gonio_cylinder_light(i)) for i in range(81)]
    g = RayTracerScene(d.width, d.height, d.vpi)
    d.populate(g, gpu_material)
    g.commitScene()
    mean, imgs, st = g.renderRequests(reqs, use_graph=True, images=True)
    assert st["paths"] == 81 * 90 * 90 * 4
    base = 0
    for k, (q, img) in enumerate(zip(reqs, imgs)):
        if k % 4 != 0 and k != 80:
            base += 90 * 90
            continue
        d2 = copy.copy(d)
        d2.lights = [dict(d.lights[0], direction=tuple(q["light_direction"]))]
        o = oracle.OracleScene(90, 90, d.vpi)
        d2.populate(o, oracle.material)
        o.setPixelKeyBase(base)
        o.commitScene()
        rgb, _ = o.renderPaths(np.arange(90 * 90, dtype=np.uint32), frame=0)
        ref = rgb.reshape(90 * 90, 4, 3).astype(np.float64).mean(axis=1).reshape(90, 90, 3)
        tol = 1e-4 * np.maximum(np.maximum(np.abs(img), np.abs(ref)), 1e-6)
        assert not (np.abs(img - ref) > tol).any(), (k, np.abs(img - ref).max())
        base += 90 * 90


def test_ray_query_matches_oracle():
    print("case: rayQuery", flush=True)
    desc = workloads.shapes(40, 40)
    g, o = _pair(desc)
    n_hit = 0
    for ty in np.linspace(0.03, 0.97, 9):
        for tx in np.linspace(0.03, 0.97, 9):
            gid, sd = g.rayQuery((tx, ty))
            oid, out = o.rayQuery((tx, ty))
            assert gid == oid, (tx, ty, gid, oid)
            if sd is None:
                continue
            n_hit += 1
            got = np.array([sd.t] + list(sd.normal) + list(sd.tangent) + list(sd.binormal) + list(sd.texcoord),
                           dtype=np.float32)
            assert np.array_equal(got.view(np.uint32), out.view(np.uint32)), (tx, ty, got, out)
    assert n_hit > 40
    # a blocker's material is NULL, a shape's the pointer it was added with
    g2 = RayTracerScene(8, 8, workloads.perspective_view_proj_inv((0, 0, 3), (0, 0, 0), (0, 1, 0), 30.0, 1.0))
    g2.addBlocker((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), (0.2, 0.2))
    g2.commitScene()
    gid, sd = g2.rayQuery((0.5, 0.5))
    assert gid == 0 and sd.material is None and sd.t == pytest.approx(2.9, rel=1e-4)   # from the near plane
    gid, sd = g2.rayQuery((0.0, 0.0))
    assert gid == 0xFFFFFFFF and sd is None


@pytest.mark.parametrize("mode", [1, 2, 3, 4])
def test_render_modes(mode):
    print("case: render mode", mode, flush=True)
    desc = workloads.shapes(32, 32, samples_camera=2, max_depth=4)
    g, o = _pair(desc)
    g.setRenderMode(mode)
    o.setRenderMode(mode)
    px = np.arange(desc.pixels, dtype=np.uint32)
    g_rgb, g_hash = g.renderPaths(px, 2)
    o_rgb, o_hash = o.renderPaths(px)
    _check_paths(g_rgb, o_rgb, g_hash, o_hash)
    assert np.all(g_rgb >= 0.0) and np.all(g_rgb <= 1.0)


def test_render_mode_arguments():
    g = RayTracerScene(8, 8, workloads.box(8, 8).vpi)
    with pytest.raises(TempestError):
        g.setRenderMode(5)
    d = workloads.volume(8, 8, samples_camera=1, max_depth=1, grid=8)
    gv = RayTracerScene(d.width, d.height, d.vpi)
    d.populate(gv, gpu_material)
    gv.commitScene()
    gv.setRenderMode(3)
    with pytest.raises(TempestError):
        gv.drawOnce()


@pytest.mark.parametrize("post", [1, 2])
def test_transparent_background(post):
    print("case: transparent background post %d" % post, flush=True)
    desc = workloads.shapes(48, 48, samples_camera=4, max_depth=3)
    desc.vpi = workloads.perspective_view_proj_inv((0.3, 0.2, 6.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, 1.0)
    # without the directional light: it reaches the back of the one-sided
    # ceiling with a negative cosine (no cosine guard in the reference's
    # BRDFs), and a negative mean pixel has no logarithm for the exposure
    desc.lights = []
    g, o = _pair(desc)
    g.setTransparentBackground(True)
    g.setPicturePostProcess(post)
    prev = np.zeros(desc.pixels, dtype=np.uint32)
    for frame in range(2):
        g.drawOnce()
        acc, spp = g.accumulation()
        mask = o.cameraHits(frame)
        assert 0 < mask.sum() < mask.size
        e_gpu = g.stats()["exposure"]
        want, e_orc = oracle.resolve(acc.reshape(-1, 3), 1.0 / spp, post, exposure=e_gpu, mask=mask, initial=prev)
        got = g.backbuffer().reshape(-1)
        assert np.array_equal(got, want), "%d of %d pixels differ" % (np.sum(got != want), got.size)
        if frame == 0:
            assert np.all(got[mask == 0] == 0) and np.all(got[mask == 1] >> 24 == 255)
        prev = got
        if post == 2:
            assert e_gpu == pytest.approx(e_orc, rel=1e-6)
