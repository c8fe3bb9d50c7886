"""The anisotropic GGX and Beckmann models of the oracle (models 15-17 and
31-33: ggx-models-impl.hh:111-127, 283-432; beckmann-models-impl.hh:69-85,
174-398), pinned by what the mathematics fixes:

- with equal roughness on both axes every variant (Schlick, conductor,
  dielectric) reduces to the isotropic model of the same family, BRDF and
  PDF, at random directions (pins the Fresnel choice, the G1 form and the
  shared factors);
- the distribution against the textbook slope-space form of the
  anisotropic NDFs (Heitz 2014, "Understanding the masking-shadowing
  function", eq. 9 for Beckmann and 13 for GGX), written in the micro-normal's
  components (m.T, m.B, m.N) rather than the reference's cos/tan terms: a
  swapped axis or alpha fails it;
- the NDF's mass below 45 degrees at normal incidence: the quadrature of the
  specular PDF over the reflected hemisphere equals
  (1 / (2 pi ax ay)) * integral over phi of 1 / (1 + a(phi)) for GGX and
  (1 / (pi ax ay)) * integral of (1 - exp(-a)) / (2 a) for Beckmann, with
  a(phi) = cos^2/ax^2 + sin^2/ay^2, from integrating the slope densities over
  the unit disc;
- shadowing: both G1 terms take the roughness projected on the outgoing
  direction, so with the outgoing direction in the tangent-normal
  (binormal-normal) plane BRDF / PDF (in which D cancels) equals the
  isotropic model's at alpha = ax (ay);
- the sampler: slopes of the sampled micro-normal have the known marginals
  (Beckmann: Gaussian, E[s^2] = alpha^2 / 2 per axis; GGX: median |s| =
  alpha / sqrt(3) per axis), measured in makeBasis(Normal), the frame the
  reference samples in.
"""
import math

import numpy as np
import pytest

import _oracle_units as U
import oracle

GGX_ANISO = (15, 16, 17)
BECK_ANISO = (31, 32, 33)
T = (1.0, 0.0, 0.0)
B = (0.0, 1.0, 0.0)
N = (0.0, 0.0, 1.0)


def _power(alpha):
    return 2.0 / (alpha * alpha) - 2.0


def _mat(model, ax, ay, diffuse=(0, 0, 0), specular=(1, 1, 1), fresnel=(1.0, 0.0)):
    return oracle.material(model, dict(diffuse=diffuse, specular=specular, specular_power=(_power(ax), _power(ay)),
                                       fresnel=fresnel))


def _unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def _random_upper(rng, n):
    d = rng.normal(size=(n, 3))
    d[:, 2] = np.abs(d[:, 2]) + 0.05
    return d / np.linalg.norm(d, axis=1, keepdims=True)


@pytest.mark.parametrize("model", GGX_ANISO + BECK_ANISO)
def test_equal_alphas_reduce_to_isotropic(model):
    iso = model - 4 if model in GGX_ANISO else model - 3
    fres = (1.5, 0.0) if model in (17, 33) else ((0.2, 3.0) if model in (16, 32) else (0.04, 0.0))
    rng = np.random.default_rng(model)
    wo = _random_upper(rng, 64)
    wi = _random_upper(rng, 64)
    for alpha in (0.15, 0.5):
        a = _mat(model, alpha, alpha, diffuse=(0.2, 0.3, 0.1), specular=(0.6, 0.5, 0.7), fresnel=fres)
        b = _mat(iso, alpha, alpha, diffuse=(0.2, 0.3, 0.1), specular=(0.6, 0.5, 0.7), fresnel=fres)
        for o, i in zip(wo, wi):
            fa, pa = U.bsdf_eval(a, N, o, i, t=T, b=B)
            fb, pb = U.bsdf_eval(b, N, o, i, t=T, b=B)
            np.testing.assert_allclose(fa, fb, rtol=3e-5, atol=1e-7)
            assert pa == pytest.approx(pb, rel=3e-5, abs=1e-7)


def _heitz_d(beckmann, ax, ay, m):
    """Anisotropic NDF in the micro-normal's frame components."""
    mx, my, mz = m
    if mz <= 0.0:
        return 0.0
    q = (mx / ax) ** 2 + (my / ay) ** 2
    if beckmann:
        return math.exp(-q / (mz * mz)) / (math.pi * ax * ay * mz ** 4)
    return 1.0 / (math.pi * ax * ay * (q + mz * mz) ** 2)


@pytest.mark.parametrize("beckmann", [False, True])
def test_distribution_matches_slope_space_form(beckmann):
    model = 31 if beckmann else 15
    ax, ay = 0.2, 0.55
    mat = _mat(model, ax, ay)
    rng = np.random.default_rng(3)
    # a rotated tangent frame, so the test also fixes which axis is which
    n = _unit((0.3, -0.2, 1.0))
    t = _unit(np.cross(n, (0.4, 1.0, 0.1)))
    b = np.cross(n, t)
    for _ in range(200):
        lo = np.array([rng.normal(), rng.normal(), abs(rng.normal()) + 0.3])
        wo = _unit(lo[0] * t + lo[1] * b + lo[2] * n)
        lh = np.array([rng.normal() * 0.3, rng.normal() * 0.3, 1.0])
        h = _unit(lh[0] * t + lh[1] * b + lh[2] * n)
        wi = 2.0 * np.dot(wo, h) * h - wo
        if np.dot(wi, n) <= 0.05 or np.dot(wo, h) <= 0.05:
            continue
        _, pdf = U.bsdf_eval(mat, n, wo, wi, t=t, b=b)
        # the half vector of the float32 directions the oracle receives
        wo32 = wo.astype(np.float32).astype(np.float64)
        h = _unit(wi.astype(np.float32).astype(np.float64) + wo32)
        m = (np.dot(h, t), np.dot(h, b), np.dot(h, n))
        want = _heitz_d(beckmann, ax, ay, m) * m[2] / (4.0 * np.dot(h, wo32))
        # the reference mixes (pdf0 - pdf1) * ratio + pdf1 with pdf1 =
        # cos / pi even at ratio 1, which leaves an absolute error of an ulp
        # of pdf1 in float
        assert pdf == pytest.approx(want, rel=5e-4, abs=1e-7)


def _ndf_mass_45(beckmann, ax, ay):
    phi = (np.arange(20000) + 0.5) * (2.0 * math.pi / 20000)
    a = np.cos(phi) ** 2 / ax ** 2 + np.sin(phi) ** 2 / ay ** 2
    dphi = 2.0 * math.pi / 20000
    if beckmann:
        return float(np.sum((1.0 - np.exp(-a)) / (2.0 * a)) * dphi / (math.pi * ax * ay))
    return float(np.sum(1.0 / (1.0 + a)) * dphi / (2.0 * math.pi * ax * ay))


@pytest.mark.parametrize("beckmann", [False, True])
def test_ndf_mass_at_normal_incidence(beckmann):
    ax, ay = 0.25, 0.6
    mat = _mat(31 if beckmann else 15, ax, ay)
    n_t, n_p = 160, 320
    th = (np.arange(n_t) + 0.5) * (0.5 * math.pi / n_t)
    ph = (np.arange(n_p) + 0.5) * (2.0 * math.pi / n_p)
    total = 0.0
    for t_ in th:
        w = math.sin(t_) * (0.5 * math.pi / n_t) * (2.0 * math.pi / n_p)
        for p_ in ph:
            d = (math.sin(t_) * math.cos(p_), math.sin(t_) * math.sin(p_), math.cos(t_))
            _, pdf = U.bsdf_eval(mat, N, N, d, t=T, b=B)
            total += pdf * w
    assert total == pytest.approx(_ndf_mass_45(beckmann, ax, ay), abs=2e-3)


@pytest.mark.parametrize("beckmann", [False, True])
def test_shadowing_uses_roughness_along_outgoing_direction(beckmann):
    ax, ay = 0.2, 0.6
    aniso = _mat(33 if beckmann else 17, ax, ay, fresnel=(1.5, 0.0))
    rng = np.random.default_rng(11)
    for axis, alpha in ((0, ax), (1, ay)):
        iso = _mat(30 if beckmann else 13, alpha, alpha, fresnel=(1.5, 0.0))
        for _ in range(40):
            th = rng.uniform(0.2, 1.3)
            s = rng.choice((-1.0, 1.0))
            wo = [0.0, 0.0, math.cos(th)]
            wo[axis] = s * math.sin(th)
            wi = _random_upper(rng, 1)[0]
            fa, pa = U.bsdf_eval(aniso, N, wo, wi, t=T, b=B)
            fb, pb = U.bsdf_eval(iso, N, wo, wi, t=T, b=B)
            # skip the far tails, where the PDF's (pdf0 - pdf1) + pdf1 form
            # leaves only a few significant bits
            if pa <= 1e-3 or pb <= 1e-3:
                continue
            np.testing.assert_allclose(fa / pa, fb / pb, rtol=5e-4)


def _slopes(model, ax, ay, k=160):
    mat = _mat(model, ax, ay)
    basis = np.zeros(9, dtype=np.float32)
    U.lib().orc_make_basis(U.fp(np.array(N, dtype=np.float32)), U.fp(basis))
    c0, c1 = basis[0:3].astype(np.float64), basis[3:6].astype(np.float64)
    sx, sy = [], []
    for i in range(k):
        for j in range(k):
            xi = (0.5, 0.5, 0.0, (i + 0.5) / k, (j + 0.5) / k)
            wi, pdf = U.bsdf_sample(mat, N, N, xi, t=T, b=B)
            h = _unit(np.asarray(wi, dtype=np.float64) + N)
            sx.append(-np.dot(h, c0) / h[2])
            sy.append(-np.dot(h, c1) / h[2])
    return np.array(sx), np.array(sy)


def test_beckmann_sampler_slope_variances():
    ax, ay = 0.3, 0.7
    sx, sy = _slopes(31, ax, ay)
    assert np.mean(sx ** 2) == pytest.approx(ax * ax / 2.0, rel=0.02)
    assert np.mean(sy ** 2) == pytest.approx(ay * ay / 2.0, rel=0.02)


def test_ggx_sampler_slope_medians():
    ax, ay = 0.3, 0.7
    sx, sy = _slopes(15, ax, ay)
    assert np.median(np.abs(sx)) == pytest.approx(ax / math.sqrt(3.0), rel=0.02)
    assert np.median(np.abs(sy)) == pytest.approx(ay / math.sqrt(3.0), rel=0.02)


def test_sampled_pdf_matches_pdf_function():
    rng = np.random.default_rng(5)
    for model in GGX_ANISO + BECK_ANISO:
        mat = _mat(model, 0.2, 0.5, diffuse=(0.3, 0.3, 0.3), specular=(0.5, 0.5, 0.5), fresnel=(1.5, 0.0))
        for _ in range(50):
            wo = _random_upper(rng, 1)[0]
            xi = tuple(float(v) for v in rng.uniform(size=5))
            wi, pdf = U.bsdf_sample(mat, N, wo, xi, t=T, b=B)
            _, p2 = U.bsdf_eval(mat, N, wo, wi, t=T, b=B)
            assert pdf == pytest.approx(p2, rel=1e-6, abs=1e-7)
