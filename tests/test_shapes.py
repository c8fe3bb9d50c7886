"""The oracle's user shapes (ray-tracing.cc:892-1046, intersect.hh:105-221)
against closed forms, and the GPU library's shape tests (compiled for the
host, through its debug traversal) against the oracle's brute-force loop.

Closed forms: the distance and normal where a ray meets a cylinder's side
(sqrt(r^2 - y^2) from the axis) and caps, an oblique cylinder whose axis
leans by tilt * z / half height, an axis-scaled and a rotated ellipsoid,
a rect's (u, v), an annulus' inner and outer rims; an oblique cylinder with
tilt 0 and an ellipsoid with unit scale equal the cylinder and the sphere.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle
from tempestrenderer_amd import workloads

F = ctypes.c_float
PF = ctypes.POINTER(F)
RECT, DISK, CYL, OBL, ELL = 1, 2, 3, 4, 5


def _hit(kind, params, org, d):
    lib = oracle.load()
    lib.orc_shape_intersect.restype = ctypes.c_int
    lib.orc_shape_intersect.argtypes = [ctypes.c_uint32, PF, PF, PF, PF]
    p = np.zeros(20, dtype=np.float32)
    p[:len(params)] = params
    o = np.asarray(org, dtype=np.float32)
    dd = np.asarray(d, dtype=np.float32)
    dd = (dd / np.linalg.norm(dd)).astype(np.float32)
    out = np.zeros(6, dtype=np.float32)
    h = lib.orc_shape_intersect(kind, p.ctypes.data_as(PF), o.ctypes.data_as(PF), dd.ctypes.data_as(PF),
                                out.ctypes.data_as(PF))
    return bool(h), float(out[0]), float(out[1]), float(out[2]), out[3:6].astype(np.float64)


def test_cylinder_side_and_caps():
    c = [0.0, 0.0, 0.0, 1.0, 2.0]          # centre, radius 1, half height 2
    h, t, _, _, n = _hit(CYL, c, (5.0, 0.3, 0.2), (-1.0, 0.0, 0.0))
    assert h and t == pytest.approx(5.0 - math.sqrt(1.0 - 0.09), abs=2e-6)
    assert np.allclose(n, (math.sqrt(0.91), 0.3, 0.0), atol=2e-6)
    # caps: a ray exactly along the axis has a = 0 in the side quadratic and
    # the reference's formula yields a NaN distance (no range check rejects
    # it there; the traversal's (tnear, tfar) check does); its side test
    # also rejects every ray whose projection moves away from the axis
    # (b < 0), caps included.  Slightly slanted rays towards the axis meet
    # the caps at the expected distance.
    h, t, _, _, n = _hit(CYL, c, (0.2, 0.1, 5.0), (0.0, 0.0, -1.0))
    assert not math.isfinite(t)
    assert not _hit(CYL, c, (0.3, 0.0, 5.0), (0.01, 0.0, -1.0))[0]
    d = np.array([-0.01, 0.0, -1.0])
    h, t, _, _, n = _hit(CYL, c, (0.3, 0.0, 5.0), d)
    assert h and t == pytest.approx(3.0 * np.linalg.norm(d), abs=2e-6) and np.allclose(n, (0, 0, 1))
    h, t, _, _, n = _hit(CYL, c, (0.3, 0.0, -5.0), (-0.01, 0.0, 1.0))
    assert h and t == pytest.approx(3.0 * np.linalg.norm(d), abs=2e-6) and np.allclose(n, (0, 0, -1))
    # a slanted ray through the top cap at a known point
    d = np.array([-1.0, 0.0, -1.0])
    h, t, _, _, n = _hit(CYL, c, (3.5, 0.0, 5.0), d)
    assert h and np.allclose(n, (0, 0, 1)) and t == pytest.approx(3.0 * math.sqrt(2.0), abs=1e-5)
    assert not _hit(CYL, c, (5.0, 0.0, 2.5), (-1.0, 0.0, 0.0))[0]        # above the top
    assert not _hit(CYL, c, (5.0, 1.5, 0.0), (-1.0, 0.0, 0.0))[0]        # beside
    assert not _hit(CYL, c, (5.0, 0.0, 0.0), (1.0, 0.0, 0.0))[0]         # pointing away


def test_oblique_cylinder():
    # axis leans by bending (0.5, 0): at height z it sits at x = 0.5 z
    ob = [0.0, 0.0, 0.0, 0.3, 1.0, 0.5, 0.0]
    h, t, _, _, n = _hit(OBL, ob, (5.0, 0.0, 0.8), (-1.0, 0.0, 0.0))
    assert h and t == pytest.approx(5.0 - (0.4 + 0.3), abs=2e-6)
    assert np.allclose(n, (1.0, 0.0, 0.0), atol=1e-6)
    h, t, _, _, n = _hit(OBL, ob, (-5.0, 0.0, -0.8), (1.0, 0.0, 0.0))
    assert h and t == pytest.approx(5.0 - 0.4 - 0.3, abs=2e-6)
    # bending 0 is the upright cylinder bit for bit
    rng = np.random.Generator(np.random.PCG64(4))
    c = [0.1, -0.2, 0.3, 0.7, 1.3]
    for _ in range(2000):
        o = rng.uniform(-3, 3, 3)
        d = rng.normal(size=3)
        a = _hit(CYL, c, o, d)
        b = _hit(OBL, c + [0.0, 0.0], o, d)
        assert a[0] == b[0] and (not a[0] or (a[1] == b[1] and np.array_equal(a[4], b[4])))


def _quat_axis(axis, angle):
    axis = np.asarray(axis, dtype=np.float64) / np.linalg.norm(axis)
    return list(axis * math.sin(angle / 2.0)) + [math.cos(angle / 2.0)]


def _ell(center, scale, q):
    q = np.asarray(q, dtype=np.float32)
    tn = [2 * (q[0] * q[2] + q[1] * q[3]), 2 * (q[1] * q[2] - q[0] * q[3]), 1 - 2 * (q[0] * q[0] + q[1] * q[1])]
    return list(center) + list(scale) + list(q) + [-q[0], -q[1], -q[2], q[3]] + tn


def test_ellipsoid():
    e = _ell((0, 0, 0), (2.0, 1.0, 0.5), (0, 0, 0, 1))
    h, t, _, _, n = _hit(ELL, e, (5.0, 0.0, 0.0), (-1.0, 0.0, 0.0))
    assert h and t == pytest.approx(3.0, abs=1e-6) and np.allclose(n, (1, 0, 0), atol=1e-6)
    h, t, _, _, n = _hit(ELL, e, (0.0, 0.0, 5.0), (0.0, 0.0, -1.0))
    assert h and t == pytest.approx(4.5, abs=1e-6) and np.allclose(n, (0, 0, 1), atol=1e-6)
    # turned 90 degrees about z, the long axis lies along y
    r = _ell((0, 0, 0), (2.0, 1.0, 0.5), _quat_axis((0, 0, 1), math.pi / 2))
    h, t, _, _, n = _hit(ELL, r, (0.0, 5.0, 0.0), (0.0, -1.0, 0.0))
    assert h and t == pytest.approx(3.0, abs=2e-6)
    h, t, _, _, n = _hit(ELL, r, (5.0, 0.0, 0.0), (-1.0, 0.0, 0.0))
    assert h and t == pytest.approx(4.0, abs=2e-6)
    # unit scale: the sphere of IntersectSphere
    s = _ell((0.3, -0.1, 0.2), (0.8, 0.8, 0.8), (0, 0, 0, 1))
    rng = np.random.Generator(np.random.PCG64(5))
    for _ in range(500):
        o = rng.uniform(-3, 3, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        h, t, _, _, n = _hit(ELL, s, o, d)
        dc = np.array([0.3, -0.1, 0.2]) - o
        b = dc @ d
        disc = b * b - (dc @ dc - 0.64)
        want = disc >= 0 and b >= 0
        if want and h:
            sq = math.sqrt(disc)
            tt = b - sq if b > sq else b + sq
            assert t == pytest.approx(tt, abs=2e-5)
        elif abs(disc) > 1e-4 and abs(b) > 1e-4:
            # the sphere form also rejects exiting hits, as the ellipsoid does
            assert h == (want and (b > math.sqrt(disc)))


def test_rect_and_disk():
    # a rect facing +z with half sizes (2, 1): columns (x, y, z)
    rect = [0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 2.0, 1.0, 0, 0, 1, 1]
    h, t, u, v, n = _hit(RECT, rect, (0.5, 0.25, 3.0), (0.0, 0.0, -1.0))
    assert h and t == pytest.approx(3.0) and u == pytest.approx(0.625) and v == pytest.approx(0.625)
    assert np.allclose(n, (0, 0, 1))
    assert not _hit(RECT, rect, (0.5, 0.25, -3.0), (0.0, 0.0, 1.0))[0]   # back face
    assert not _hit(RECT, rect, (2.5, 0.0, 3.0), (0.0, 0.0, -1.0))[0]    # outside
    disk = [0, 0, 0, 0, 0, 1, 0.5, 1.0]
    for r, want in ((0.3, False), (0.7, True), (0.99, True), (1.2, False)):
        h, t, _, _, n = _hit(DISK, disk, (r, 0.0, 2.0), (0.0, 0.0, -1.0))
        assert h == want
        if want:
            assert t == pytest.approx(2.0)


def test_rect_orientation_is_the_input_frame():
    """addRect's frame after the quaternion round trip (MatrixToQuaternion
    then ToMatrix3) is the given (tan, norm x tan, norm): rays hit where the
    frame puts the rect, with (u, v) along tan and norm x tan."""
    rng = np.random.Generator(np.random.PCG64(6))
    for _ in range(20):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        x, y, z, w = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        tan, norm = R[:, 0], R[:, 2]
        d = workloads.SceneDesc("r", 4, 4, workloads.IDENTITY)
        m = dict(model=workloads.GGX, diffuse=(0.5, 0.5, 0.5), specular=(0.0, 0.0, 0.0), specular_power=(1.0, 1.0))
        d.shapes.append(("addRect", dict(pos=(0.0, 0.0, 0.0), tan=tuple(tan), norm=tuple(norm), size=(1.5, 0.5)), m))
        o = oracle.OracleScene(4, 4, workloads.IDENTITY)
        d.populate(o, oracle.material)
        o.commitScene()
        a, b = 0.6, -0.3
        p = a * tan + b * np.cross(norm, tan)
        ray = np.zeros((1, 8), dtype=np.float32)
        ray[0, :3] = p + 2.0 * norm
        ray[0, 4:7] = -norm
        ray[0, 7] = np.inf
        hit = o.traceRays(ray, brute_force=True)[0]
        assert hit["geom"] == 0 and hit["t"] == pytest.approx(2.0, abs=1e-5)
        assert hit["u"] == pytest.approx(a / 1.5 * 0.5 + 0.5, abs=1e-5)
        assert hit["v"] == pytest.approx(b / 0.5 * 0.5 + 0.5, abs=1e-5)


@pytest.mark.parametrize("name", ["shapes", "gonio_cylinders"])
def test_library_shapes_match_oracle_on_host(name):
    """The GPU library's traversal and shape tests compiled for the host
    (tr_debug_trace_rays_host) give the oracle's brute-force hits bit for bit."""
    from tempestrenderer_amd.scene import Material, RayTracerScene
    d = workloads.shapes(40, 40) if name == "shapes" else workloads.gonio_cylinders(30, 30, samples_camera=4)
    o = oracle.OracleScene(d.width, d.height, d.vpi)
    d.populate(o, oracle.material)
    o.commitScene()
    rays = np.concatenate([workloads.primary_rays(d), workloads.random_rays(d, 1 << 14)])
    hb = o.traceRays(rays, brute_force=True)
    assert np.array_equal(o.traceRays(rays)["geom"], hb["geom"])
    g = RayTracerScene(d.width, d.height, d.vpi)
    d.populate(g, lambda m, f: Material(m, **f).setup())
    g.buildHostOnly()
    assert g.bvhViolations() == 0
    hh, _, _ = g.traceRaysHost(rays)
    hit = hb["geom"] != 0xFFFFFFFF
    assert np.array_equal(hh["geom"], hb["geom"]) and np.array_equal(hh["prim"], hb["prim"])
    assert np.array_equal(hh["t"][hit].view(np.uint32), hb["t"][hit].view(np.uint32))
    assert np.array_equal(hh["u"][hit], hb["u"][hit]) and np.array_equal(hh["v"][hit], hb["v"][hit])
    g.close()


def test_shape_arguments_are_checked():
    from tempestrenderer_amd._abi import TempestError
    from tempestrenderer_amd.scene import Material, RayTracerScene
    g = RayTracerScene(8, 8, workloads.IDENTITY)
    m = Material(workloads.GGX, diffuse=(0.5, 0.5, 0.5), specular=(0.0, 0.0, 0.0), specular_power=(1.0, 1.0)).setup()
    for call in (lambda: g.addCylinder((0, 0, 0), -1.0, 1.0, m),
                 lambda: g.addDisk((0, 0, 0), (0, 0, 1), 1.0, 0.5, m),
                 lambda: g.addRect((0, 0, 0), (1, 0, 0), (0, 0, 1), (0.0, 1.0), m),
                 lambda: g.addEllipsoid((0, 0, 0), (1, 1, 1), (0, 0, 0, 0), m),
                 lambda: g.addRect((0, 0, 0), (1, 0, 0), (0, 0, 1), (1.0, 1.0), m, tangent_map=3)):
        with pytest.raises(TempestError):
            call()
    g.close()
