"""Batched render requests (tr_scene_render_requests) on the GPU.

A request is a frame of its own orthographic camera, size and camera
sample count with light 0 turned to its direction (the reference's serial
loop, samples/virtual-gonioreflectometer.cc:161-189).  Each request's
per-pixel mean is compared with the oracle rendering the same request as a
frame whose random-number pixel key starts at the request's first pixel
over all requests, to the per-path contract (1e-4 relative); the batching
(graph or plain launches, pool size, autotuning) must not change a bit.
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

# (light dir index, view dir index, width, height, camera samples)
CASES = [(0, 40, 12, 8, 3), (40, 40, 16, 16, 2), (80, 10, 9, 11, 4), (20, 60, 8, 8, 1), (44, 44, 10, 6, 5),
         (72, 3, 7, 9, 2)]


def _scene_and_requests():
    d, _ = workloads.gonio()
    dirs = workloads.hemisphere_grid(9)
    reqs = [dict(view_proj_inv=workloads.ortho_view_proj_inv(dirs[v], 1.5), width=w, height=h, samples_camera=s,
                 light_direction=dirs[li]) for li, v, w, h, s in CASES]
    return d, reqs


def _gpu_scene(d):
    g = RayTracerScene(32, 32, d.vpi)
    d.populate(g, gpu_material)
    g.commitScene()
    return g


@pytest.mark.parametrize("gi,estimator", [(1, 0), (4, 1)])
def test_requests_match_oracle_per_pixel(gi, estimator):
    print("case: render requests vs oracle, 6 requests, pool 512, gi %d estimator %d" % (gi, estimator), flush=True)
    d, reqs = _scene_and_requests()
    d.settings.update(samples_gi=gi, estimator=estimator)
    g = _gpu_scene(d)
    mean, imgs, st = g.renderRequests(reqs, use_graph=True, pool_slots=512, images=True)
    total = sum(q["width"] * q["height"] * q["samples_camera"] for q in reqs) * gi
    assert st["paths"] == total and st["batches"] >= -(-total // 512) > 1   # batches cut through requests
    base = 0
    for q, img in zip(reqs, imgs):
        d2 = copy.copy(d)
        d2.lights = [dict(d.lights[0], direction=tuple(q["light_direction"]))]
        o = oracle.OracleScene(q["width"], q["height"], q["view_proj_inv"])
        d2.populate(o, oracle.material)
        o.setSamplesCamera(q["samples_camera"])
        o.setPixelKeyBase(base)
        o.commitScene()
        n = q["width"] * q["height"]
        rgb, _ = o.renderPaths(np.arange(n, dtype=np.uint32), frame=0)
        ref = rgb.reshape(n, q["samples_camera"], 3).astype(np.float64).mean(axis=1).reshape(q["height"], q["width"], 3)
        tol = 1e-4 * np.maximum(np.maximum(np.abs(img), np.abs(ref)), 1e-6)
        assert not (np.abs(img - ref) > tol).any(), (q["width"], q["height"], np.abs(img - ref).max())
        assert ref.max() > 0.0
        base += n
    for m, img in zip(mean, imgs):
        np.testing.assert_allclose(m, img.reshape(-1, 3).astype(np.float64).mean(axis=0), rtol=1e-5, atol=1e-9)


def test_requests_do_not_depend_on_batching():
    print("case: render requests graph / plain / pool sizes / autotune", flush=True)
    d, reqs = _scene_and_requests()
    g = _gpu_scene(d)
    ref_mean, ref_imgs, _ = g.renderRequests(reqs, use_graph=True, pool_slots=512, images=True)
    for kw in (dict(use_graph=False, pool_slots=512), dict(use_graph=True, pool_slots=300, trace_waves_per_cu=8),
               dict(use_graph=False, pool_slots=1 << 16, trace_waves_per_cu=32), dict(use_graph=True, autotune=True)):
        mean, imgs, st = g.renderRequests(reqs, images=True, **kw)
        assert np.array_equal(mean, ref_mean), kw
        for a, b in zip(imgs, ref_imgs):
            assert np.array_equal(a, b), kw
    # a plain frame still renders after request mode
    g.drawOnce()


def test_requests_reject_bad_arguments():
    print("case: render requests argument checks", flush=True)
    d, reqs = _scene_and_requests()
    g = _gpu_scene(d)
    bad = [dict(reqs[0], width=1), dict(reqs[0], samples_camera=0)]
    for q in bad:
        with pytest.raises(TempestError):
            g.renderRequests([q])
    d2 = copy.copy(d)
    d2.lights = []
    with pytest.raises(TempestError):
        _gpu_scene(d2).renderRequests(reqs[:1])


def test_full_gonio_workload_sampled_requests_match_oracle():
    """The whole batched gonioreflectometer workload of the benchmark (6,561
    requests, 32^2-256^2 pixels, 4-64 camera samples, 3.56 G paths) in one
    call with autotuning and graph launches; the per-request means of five
    requests spread over the call (first, last, the largest images among
    them) against the oracle rendering each request on its own with its
    pixel key base, to the per-path contract (1e-4 relative)."""
    print("case: full gonioreflectometer workload, 6561 requests, autotune + graph, 5 sampled requests", flush=True)
    d, reqs = workloads.gonio()
    g = _gpu_scene(d)
    mean, _, st = g.renderRequests(reqs, use_graph=True, autotune=True, images=False)
    g.close()
    sizes = [q["width"] * q["height"] for q in reqs]
    assert st["paths"] == sum(n * q["samples_camera"] for n, q in zip(sizes, reqs))
    base = np.concatenate([[0], np.cumsum(sizes)])
    big = [i for i in range(len(reqs)) if reqs[i]["width"] == 256]
    picks = sorted({0, len(reqs) - 1, len(reqs) // 2, big[0], big[-1]})
    for r in picks:
        q = reqs[r]
        d2 = copy.copy(d)
        d2.lights = [dict(d.lights[0], direction=tuple(q["light_direction"]))]
        o = oracle.OracleScene(q["width"], q["height"], q["view_proj_inv"])
        d2.populate(o, oracle.material)
        o.setSamplesCamera(q["samples_camera"])
        o.setPixelKeyBase(int(base[r]))
        o.commitScene()
        n = sizes[r]
        rgb, _ = o.renderPaths(np.arange(n, dtype=np.uint32), frame=0)
        ref = rgb.reshape(n, q["samples_camera"], 3).astype(np.float64).mean(axis=1).mean(axis=0)
        print("request %d (%dx%d, %d spp): GPU %s oracle %s" % (r, q["width"], q["height"], q["samples_camera"],
                                                                mean[r], ref), flush=True)
        assert ref.max() > 0.0
        np.testing.assert_allclose(mean[r], ref, rtol=1e-4, atol=1e-9)
