"""Closed forms that pin the oracle's path loop (restating
ray-tracing.cc:1981-2379) independently of its own formulas.

- Furnace: a Lambertian plane of albedo rho under a constant environment L:
  every bounce escapes, each path carries exactly rho * L.
- Direct light from a sphere light (depth 0, NEE only, no MIS): a point
  directly under a sphere of radius R at height h reflects
  rho * L * (R/h)^2 (irradiance pi L sin^2(alpha) of a spherical source).
- Camera rays that miss return the background; camera rays that hit an
  emitter return its radiance (ray-tracing.cc:2057-2061, 2148-2161).
"""
import math

import numpy as np

import oracle
from tempestrenderer_amd import workloads


def _oracle(desc):
    o = oracle.OracleScene(desc.width, desc.height, desc.vpi)
    desc.populate(o, oracle.material)
    o.commitScene()
    return o


def _plane_desc(albedo, width=16, height=16, depth=3, spp=4):
    vpi = workloads.perspective_view_proj_inv((0.0, 1.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, -1.0), 30.0, 1.0)
    d = workloads.SceneDesc("plane", width, height, vpi)
    qv, qt = workloads._quad((0.0, 0.0, 0.0), (0.0, 0.0, 50.0), (50.0, 0.0, 0.0))  # normal +y
    d.meshes.append((qv, qt, [(dict(model=workloads.GGX, diffuse=albedo, specular=(0, 0, 0)), 0, 6)], False))
    d.settings.update(samples_camera=spp, samples_gi=1, samples_light=1, max_depth=depth, background=(1.0, 1.0, 1.0))
    return d


def test_furnace_plane_under_constant_sky():
    d = _plane_desc((0.7, 0.5, 0.3))
    o = _oracle(d)
    rgb, _ = o.renderPaths(np.arange(d.pixels, dtype=np.uint32))
    np.testing.assert_allclose(rgb, np.tile([0.7, 0.5, 0.3], (len(rgb), 1)), rtol=3e-6)


def test_sphere_light_irradiance():
    h, R, L, rho = 2.0, 0.5, 4.0, 0.6
    d = _plane_desc((rho, rho, rho), width=1, height=1, depth=0, spp=20000)
    d.vpi = workloads.perspective_view_proj_inv((0.0, 1.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, -1.0), 0.01, 1.0)
    d.settings.update(background=(0.0, 0.0, 0.0))
    d.sphere_lights.append(dict(kind=workloads.LIGHT_SPHERE_AREA, radiance=(L, L, L), center=(0.0, h, 0.0), radius=R))
    o = _oracle(d)
    rgb, _ = o.renderPaths(np.array([0], dtype=np.uint32))
    est = rgb[:, 0].astype(np.float64)
    expect = rho * L * (R / h) ** 2
    sigma = est.std() / math.sqrt(len(est))
    assert abs(est.mean() - expect) < 4 * sigma + 1e-4 * expect, (est.mean(), expect, sigma)


def test_background_and_emitter_at_camera_vertex():
    vpi = workloads.perspective_view_proj_inv((0.0, 0.0, 3.5), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, 1.0)
    d = workloads.SceneDesc("emitter", 32, 32, vpi)
    qv, qt = workloads._quad((0.0, 0.0, 0.0), (0.4, 0.0, 0.0), (0.0, 0.4, 0.0))  # x cross y: faces the camera
    d.meshes.append((qv, qt, [(dict(model=workloads.EMISSIVE, diffuse=(15.0, 7.0, 3.0)), 0, 6)], False))
    d.settings.update(samples_camera=1, samples_gi=1, max_depth=2, background=(0.25, 0.5, 0.75))
    o = _oracle(d)
    hits = o.traceRays(workloads.primary_rays(d), brute_force=True)
    rgb, _ = o.renderPaths(np.arange(d.pixels, dtype=np.uint32))
    # with one camera sample the camera ray is the pixel's corner-to-corner
    # grid ray (ray-tracing.cc:2018-2022), not the centre ray: compare per path
    miss = np.all(rgb == np.float32([0.25, 0.5, 0.75]), axis=1)
    emit = np.all(rgb == np.float32([15.0, 7.0, 3.0]), axis=1)
    assert miss.sum() > 0 and emit.sum() > 0
    assert (miss | emit).all()
    assert abs(emit.mean() - np.mean(hits["geom"] == 0)) < 0.05


def _rect_form_factor(x, y):
    """Point-to-parallel-rectangle form factor, the normal through one
    corner of an x by y rectangle (in units of the distance)."""
    a = x / math.sqrt(1 + x * x) * math.atan(y / math.sqrt(1 + x * x))
    b = y / math.sqrt(1 + y * y) * math.atan(x / math.sqrt(1 + y * y))
    return (a + b) / (2 * math.pi)


def _direct_desc(estimator, light, rho=0.5, spp=40000):
    h = 0.5
    vpi = workloads.perspective_view_proj_inv((0.0, h / 2, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, -1.0), 0.2, 1.0)
    d = workloads.SceneDesc("direct", 1, 1, vpi)
    qv, qt = workloads._quad((0.0, 0.0, 0.0), (0.0, 0.0, 50.0), (50.0, 0.0, 0.0))  # normal +y
    plane = dict(model=workloads.GGX, diffuse=(rho, rho, rho), specular=(0, 0, 0))
    parts = [(plane, 0, 6)]
    verts, tris = [qv], [qt]
    if light == "quad":
        lv, lt = workloads._quad((0.0, h, 0.0), (0.5, 0.0, 0.0), (0.0, 0.0, 0.5))   # 1 x 1, faces -y
        parts.append((dict(model=workloads.EMISSIVE, diffuse=(3.0, 3.0, 3.0)), 6, 6))
        verts.append(lv)
        tris.append(lt + 4)
    else:
        d.sphere_lights.append(dict(kind=workloads.LIGHT_SPHERE_AREA, radiance=(3.0, 3.0, 3.0), center=(0.0, 2.0, 0.0),
                                    radius=0.5))
    d.meshes.append((np.concatenate(verts), np.concatenate(tris), parts, False))
    d.settings.update(samples_camera=spp, samples_gi=4, samples_light=2, max_depth=1, background=(0.0, 0.0, 0.0),
                      estimator=estimator)
    return d


def _mean_and_error(o):
    rgb, _ = o.renderPaths(np.array([0], dtype=np.uint32), threads=8)
    est = rgb[:, 0].astype(np.float64)
    return est.mean(), est.std() / math.sqrt(len(est))


def test_unbiased_mesh_light_direct_illumination():
    """Unbiased estimator, depth 1 (NEE and BSDF sampling both active, MIS
    with 2 light samples and 4 BSDF samples): a Lambertian point under a
    1 x 1 emitting quad at height 0.5 reflects rho * L * 4 F(1, 1).  The
    reference estimator, whose mesh-light NEE is not MIS-weighted while its
    BSDF hits are, over-counts the same light."""
    expect = 0.5 * 3.0 * 4.0 * _rect_form_factor(1.0, 1.0)
    m, e = _mean_and_error(_oracle(_direct_desc(1, "quad")))
    assert abs(m - expect) < 4.0 * e + 1e-4 * expect, (m, expect, e)
    m_ref, e_ref = _mean_and_error(_oracle(_direct_desc(0, "quad")))
    assert m_ref - expect > 6.0 * e_ref, (m_ref, expect, e_ref)


def test_unbiased_sphere_light_direct_illumination():
    """Sphere light of radius 0.5 at height 2 above the point, MIS active:
    rho * L * (R / h)^2."""
    expect = 0.5 * 3.0 * (0.5 / 2.0) ** 2
    m, e = _mean_and_error(_oracle(_direct_desc(1, "sphere")))
    assert abs(m - expect) < 4.0 * e + 1e-4 * expect, (m, expect, e)
