"""GPU parity: the HIP kernels against the CPU oracle, through the C ABI.

Contract: closest-hit (geom, prim) ids and t bit-exact on fixed ray sets;
per-path radiance within 1e-4 relative (|a-b| <= 1e-4 max(|a|,|b|,1e-6))
with identical hit-sequence hashes; frame accumulation equal to the sum of
the per-path values.
"""
import numpy as np
import pytest

import oracle
from tempestrenderer_amd import workloads
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


def _ids_equal(a, b):
    return (a["geom"] == b["geom"]) & (a["prim"] == b["prim"])


def _check_paths(g_rgb, o_rgb, g_hash, o_hash):
    assert np.array_equal(g_hash, o_hash), "hit sequences differ on %d paths" % np.sum(g_hash != o_hash)
    tol = 1e-4 * np.maximum(np.maximum(np.abs(g_rgb), np.abs(o_rgb)), 1e-6)
    bad = np.abs(g_rgb - o_rgb) > tol
    assert not bad.any(), "radiance differs on %d of %d values" % (bad.sum(), bad.size)


@pytest.mark.parametrize("scene", ["box", "heightfield256"])
def test_trace_ids_match_bruteforce(scene):
    print("case: trace ids", scene, flush=True)
    desc = workloads.box(64, 64) if scene == "box" else workloads.heightfield(160, 90, grid=256)
    g, o = _pair(desc)
    rays = np.concatenate([workloads.primary_rays(desc), workloads.random_rays(desc, 1 << 16)])
    hb = o.traceRays(rays, brute_force=True)
    hg = g.traceRays(rays)
    assert _ids_equal(hb, hg).all()
    hit = hb["geom"] != 0xFFFFFFFF
    assert np.array_equal(hb["t"][hit], hg["t"][hit])
    assert np.array_equal(hb["u"][hit], hg["u"][hit]) and np.array_equal(hb["v"][hit], hg["v"][hit])


def test_paths_box():
    print("case: paths box 64x64 spp4 depth4", flush=True)
    desc = workloads.box(64, 64, samples_camera=4, max_depth=4)
    g, o = _pair(desc)
    px = np.arange(desc.pixels, dtype=np.uint32)
    g_rgb, g_hash = g.renderPaths(px, 4)
    o_rgb, o_hash = o.renderPaths(px)
    _check_paths(g_rgb, o_rgb, g_hash, o_hash)


def test_paths_heightfield():
    print("case: paths heightfield grid 256", flush=True)
    desc = workloads.heightfield(96, 54, samples_camera=4, max_depth=8, grid=256)
    g, o = _pair(desc)
    px = np.arange(desc.pixels, dtype=np.uint32)
    g_rgb, g_hash = g.renderPaths(px, 4)
    o_rgb, o_hash = o.renderPaths(px)
    _check_paths(g_rgb, o_rgb, g_hash, o_hash)


@pytest.mark.parametrize("enclosing", [False, True])
@pytest.mark.parametrize("estimator,samples_light", [(0, 1), (0, 2), (1, 1), (1, 2)])
def test_paths_sphere_light(enclosing, estimator, samples_light):
    """A sphere-area light (addSphereLightSource): the cone sampler
    SampleSphereArea and its PDF from outside, the uniform area sampler from
    inside (enclosing sphere), the reference estimator's balance-weighted
    area-light NEE (ray-tracing.cc:1852-1860) and the unbiased estimator's
    multi-sample weights, with 1 and 2 samples per area light.  Inside the
    enclosing sphere, shadow rays that leave through the back of a one-sided
    wall reach the light with a negative cosine; the reference's BRDFs have
    no cosine guard (ggx-models-impl.hh:168-179), so such samples subtract,
    on both sides alike."""
    print("case: paths sphere light enclosing=%d estimator=%d samples_light=%d" % (enclosing, estimator,
                                                                                   samples_light), flush=True)
    desc = workloads.sphere_light_box(40, 40, samples_camera=4, max_depth=4, enclosing=enclosing)
    desc.settings.update(estimator=estimator, samples_light=samples_light)
    g, o = _pair(desc)
    px = np.arange(desc.pixels, dtype=np.uint32)
    g_rgb, g_hash = g.renderPaths(px, 4)
    o_rgb, o_hash = o.renderPaths(px)
    _check_paths(g_rgb, o_rgb, g_hash, o_hash)


def test_draw_once_accumulates_paths():
    print("case: drawOnce box", flush=True)
    desc = workloads.box(48, 48, samples_camera=4, max_depth=3)
    g, o = _pair(desc)
    g.drawOnce()
    acc, spp = g.accumulation()
    assert spp == 4
    o_sum, _, _ = o.renderRows(0, desc.height)
    np.testing.assert_allclose(acc, o_sum, rtol=1e-5, atol=1e-6)
    g.drawOnce()
    acc2, spp2 = g.accumulation()
    assert spp2 == 8
    assert np.all(np.isfinite(acc2))


def test_trace_ids_volume_boxes():
    """Fixed rays against a scene with a grid volume: the volume box is a
    primitive whose hit is the entry distance with the exit in u."""
    print("case: trace ids volume", flush=True)
    desc = workloads.volume(32, 32, samples_camera=1, max_depth=2, grid=16, with_floor_light=True)
    g, o = _pair(desc)
    rays = np.concatenate([workloads.primary_rays(desc), workloads.random_rays(desc, 1 << 14)])
    hb = o.traceRays(rays, brute_force=True)
    hg = g.traceRays(rays)
    assert _ids_equal(hb, hg).all()
    hit = hb["geom"] != 0xFFFFFFFF
    assert np.array_equal(hb["t"][hit], hg["t"][hit]) and np.array_equal(hb["u"][hit], hg["u"][hit])


def test_paths_volume():
    print("case: paths volume 48x48 spp4 depth12 grid32", flush=True)
    desc = workloads.volume(48, 48, samples_camera=4, max_depth=12, grid=32)
    g, o = _pair(desc)
    px = np.arange(desc.pixels, dtype=np.uint32)
    g_rgb, g_hash = g.renderPaths(px, 4)
    o_rgb, o_hash = o.renderPaths(px)
    _check_paths(g_rgb, o_rgb, g_hash, o_hash)


def test_paths_volume_floor_light():
    """Next-event estimation through the volume: Woodcock shadow rays."""
    print("case: paths volume + floor + directional light", flush=True)
    desc = workloads.volume(40, 40, samples_camera=4, max_depth=6, grid=24, with_floor_light=True)
    g, o = _pair(desc)
    px = np.arange(desc.pixels, dtype=np.uint32)
    g_rgb, g_hash = g.renderPaths(px, 4)
    o_rgb, o_hash = o.renderPaths(px)
    _check_paths(g_rgb, o_rgb, g_hash, o_hash)


def test_trace_ids_brick_volume():
    """The reference's brick volumes (addHierarchicalVolume): one box per
    non-empty brick, the empty brick absent."""
    print("case: trace ids brick volume", flush=True)
    desc = workloads.brick_volume(32, 32, samples_camera=1, max_depth=2)
    g, o = _pair(desc)
    rays = np.concatenate([workloads.primary_rays(desc), workloads.random_rays(desc, 1 << 14)])
    hb = o.traceRays(rays, brute_force=True)
    hg = g.traceRays(rays)
    assert _ids_equal(hb, hg).all()
    hit = hb["geom"] != 0xFFFFFFFF
    assert np.array_equal(hb["t"][hit], hg["t"][hit]) and np.array_equal(hb["u"][hit], hg["u"][hit])


def test_paths_brick_volume():
    """u8 densities, quantized fibre angles through the cos/sin tables and
    sampleScattering's brick arithmetic, with Woodcock shadow rays."""
    print("case: paths brick volume 40x40 spp4 depth8", flush=True)
    desc = workloads.brick_volume(40, 40, samples_camera=4, max_depth=8)
    g, o = _pair(desc)
    px = np.arange(desc.pixels, dtype=np.uint32)
    g_rgb, g_hash = g.renderPaths(px, 4)
    o_rgb, o_hash = o.renderPaths(px)
    _check_paths(g_rgb, o_rgb, g_hash, o_hash)


def test_paths_textured():
    """DiffuseMap on Blinn-Phong (floor, sphere), the four SGGX-surface maps
    (wall), wrapping bilinear and nearest filters, and the cube-map
    environment."""
    print("case: paths textured 48x48 spp4 depth4", flush=True)
    desc = workloads.textured(48, 48, samples_camera=4, max_depth=4)
    g, o = _pair(desc)
    px = np.arange(desc.pixels, dtype=np.uint32)
    g_rgb, g_hash = g.renderPaths(px, 4)
    o_rgb, o_hash = o.renderPaths(px)
    _check_paths(g_rgb, o_rgb, g_hash, o_hash)


def test_paths_anisotropic():
    """The six anisotropic GGX / Beckmann models on tangent-frame meshes,
    spheres and a mesh without tangents, with their second stratified pair
    in the specular lobe."""
    print("case: paths anisotropic 48x48 spp4 depth4", flush=True)
    desc = workloads.anisotropic(48, 48, samples_camera=4, max_depth=4)
    g, o = _pair(desc)
    px = np.arange(desc.pixels, dtype=np.uint32)
    g_rgb, g_hash = g.renderPaths(px, 4)
    o_rgb, o_hash = o.renderPaths(px)
    _check_paths(g_rgb, o_rgb, g_hash, o_hash)


def test_paths_pseudo_volume():
    """SGGX pseudo-volume surfaces: TexCoord-seeded random walks inside
    the BRDF, mapped and constant parameters, a channel under the walks'
    cut."""
    print("case: paths pseudo-volume 32x32 spp2 depth3", flush=True)
    desc = workloads.pseudo_volume(32, 32, samples_camera=2, max_depth=3)
    g, o = _pair(desc)
    px = np.arange(desc.pixels, dtype=np.uint32)
    g_rgb, g_hash = g.renderPaths(px, 2)
    o_rgb, o_hash = o.renderPaths(px)
    _check_paths(g_rgb, o_rgb, g_hash, o_hash)


def test_pseudo_volume_arguments_are_checked():
    print("case: pseudo-volume arguments", flush=True)
    from tempestrenderer_amd._abi import TempestError
    desc = workloads.box(8, 8)
    for bad in (dict(sggx_sample_count=0), dict(sggx_sample_count=5000), dict(sggx_depth=65)):
        g = RayTracerScene(8, 8, desc.vpi)
        g.addSphere((0, 0, 0), 0.5, gpu_material(workloads.SGGX_PSEUDO_VOLUME, bad))
        with pytest.raises(TempestError):
            g.commitScene()


def test_pseudo_volume_benchmark_image():
    """The ComputeImage microbenchmark: every pixel holds ToColor of the
    same BRDF value, which matches the oracle's at the returned directional
    density; that density is the closed-form projected area
    sqrt(0.25 sin^2 + cos^2) of stddev (0.5, 0.5, 1) at 45 degrees."""
    print("case: pseudo-volume benchmark image 64x32", flush=True)
    import ctypes
    import _oracle_units as U
    from tempestrenderer_amd import _abi
    fields = dict(diffuse=(0.0, 0.0, 0.0), specular=(0.75, 0.75, 0.75), specular_power=(0.5, 0.5),
                  sggx_basis=(0.0, 0.0, 0.0, 1.0), sggx_depth=1, sggx_sample_count=256)
    m = gpu_material(workloads.SGGX_PSEUDO_VOLUME, fields)
    img = np.zeros(64 * 32, dtype=np.uint32)
    probe = np.zeros(4, dtype=np.float32)
    ms = ctypes.c_float()
    _abi.check(_abi.load().tr_bench_pseudo_volume_image(
        ctypes.byref(m.rec), 64, 32, 2, img.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
        probe.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(ms)))
    assert ms.value > 0.0
    assert np.all(img == img[0])
    sn = float(U.fmath(0, np.array([0.125], np.float32))[0])
    cs = float(U.fmath(1, np.array([0.125], np.float32))[0])
    assert probe[3] == pytest.approx(np.sqrt(0.25 * 0.5 + 0.5), rel=1e-6)
    om = oracle.material(workloads.SGGX_PSEUDO_VOLUME, fields)
    f, _ = U.bsdf_eval(om, (0, 0, 1), (0, -sn, cs), (0, sn, cs), t=(1, 0, 0), b=(0, 1, 0),
                       dir_density=float(probe[3]))
    np.testing.assert_allclose(probe[:3], f, rtol=1e-5)
    q = [int(np.float32(255.0) * np.float32(min(max(v, 0.0), 1.0))) for v in probe[:3]]
    assert img[0] == (q[0] | (q[1] << 8) | (q[2] << 16) | (255 << 24))


def test_texture_arguments_are_checked():
    print("case: texture arguments", flush=True)
    from tempestrenderer_amd._abi import TempestError
    from tempestrenderer_amd.scene import Material, Submesh
    desc = workloads.box(8, 8)
    g = RayTracerScene(8, 8, desc.vpi)
    with pytest.raises(TempestError):
        g.bindTexture(np.zeros((4, 4), np.uint32), 9)
    with pytest.raises(TempestError):
        g.setGlobalCubeMap([np.zeros((16, 16), np.uint32)] * 6)
    rg = g.bindTexture(np.zeros((4, 4, 2), np.uint8), 1)
    qv, qt = workloads._textured_quad((0, 0, 0), (0, 0, 1), (1, 0, 0))
    g.addTriangleMesh(workloads.IDENTITY, [Submesh(Material(1, diffuse_map=rg), 6, 0, 0, 32)], qt, qv)
    with pytest.raises(TempestError):
        g.commitScene()  # an RG8 texture cannot be a DiffuseMap
    g2 = RayTracerScene(8, 8, desc.vpi)
    g2.addTriangleMesh(workloads.IDENTITY, [Submesh(Material(1, diffuse_map=3), 6, 0, 0, 32)], qt, qv)
    with pytest.raises(TempestError):
        g2.commitScene()  # no texture 3


def test_timing_mode_renders_the_same_frame():
    """Launches that carry timing events compute what plain launches do."""
    print("case: timing mode box", flush=True)
    desc = workloads.box(32, 32, samples_camera=2, max_depth=3)
    g, _ = _pair(desc)
    g.setFrameIndex(0)
    g.drawOnce()
    a, _ = g.accumulation()
    g.repaint()
    g.setFrameIndex(0)
    g.setTiming(True)
    g.drawOnce()
    st = g.stats()
    b, _ = g.accumulation()
    assert np.array_equal(a, b)
    assert st["trace_ms"] > 0.0 and st["shade_ms"] > 0.0 and st["trace_launches"] == 4


@pytest.mark.parametrize("estimator", [0, 1])
def test_paths_box_multi_sample(estimator):
    """4 GI splits and 2 samples per area light, both estimators: the
    unbiased one MIS-weights the mesh light's NEE and the emitter hits."""
    print("case: paths box multi-sample estimator", estimator, flush=True)
    desc = workloads.box(40, 40, samples_camera=2, max_depth=4)
    desc.settings.update(estimator=estimator, samples_gi=4, samples_light=2)
    g, o = _pair(desc)
    px = np.arange(desc.pixels, dtype=np.uint32)
    g_rgb, g_hash = g.renderPaths(px, 2)
    o_rgb, o_hash = o.renderPaths(px)
    _check_paths(g_rgb, o_rgb, g_hash, o_hash)


def test_unbiased_rejects_volumes():
    desc = workloads.volume(16, 16, samples_camera=1, max_depth=2, grid=8)
    desc.settings.update(estimator=1)
    g = RayTracerScene(desc.width, desc.height, desc.vpi)
    desc.populate(g, gpu_material)
    g.commitScene()
    from tempestrenderer_amd import _abi
    with pytest.raises(_abi.TempestError) as e:
        g.drawOnce()
    assert e.value.status == _abi.TR_ERROR_UNSUPPORTED
