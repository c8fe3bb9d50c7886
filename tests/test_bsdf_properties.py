"""Properties of the oracle's BSDFs that the reference's own model tests
check, with the same or tighter tolerances:

- microfacet NDF normalisation: at normal incidence the specular-lobe PDF
  D(h) cos(h) / (4 h.o) integrated over the reflected hemisphere is the
  mass of D cos with theta_h < 45 deg, which has closed forms: 1/(1+a^2)
  for GGX, 1 - exp(-1/a^2) for Beckmann (the reference's own tests check
  the full NDF integral to +-0.1 / +-0.2: ggx-test.cc:61-72,
  beckmann-test.cc:56-80; here +-2e-3);
- energy: the BRDF*cos integrates to at most 1 for a white lobe with F = 1
  (ggx-test.cc:74-119 weak furnace +-0.1; brdf-test.cc:66-85 <= 1.1), and
  close to 1 for a smooth lobe;
- sampling consistency: the PDF the sampler returns equals the PDF function
  at the sampled direction, and E[f cos / pdf] from the sampler matches the
  quadrature of f cos (importance sampling is unbiased only if the sampler
  draws from its PDF);
- SGGX: with the reference's 0.5 factor (ggx-models-impl.hh:48-57) the
  integral of D(w) |w.v| over the sphere equals the projected area
  sqrt(v^T S v) that SGGXProjectedArea returns, and the isotropic NDF is the
  constant 0.5/pi (sggx-test.cc:418-469 closed-form identities, +-1e-3);
- the Lambertian case (Specular 0) gives BRDF*cos/pdf = albedo exactly;
- Blinn-Phong (illumination-models.hh:618-698): the half-vector density
  (n+1) cos^n(h) / (2 pi) of PowerCosineSampleHemisphere carried to the
  reflected direction integrates, at normal incidence, to the mass of h
  below 45 deg, 1 - 2^(-(n+1)/2); its half-vector warp equals the
  reference's PowerCosineSampleHemisphere (tests/golden/ref_phong.npz).
"""
import os
import math

import numpy as np
import pytest

import _oracle_units as U
import oracle

PI32 = 3.141592


def _hemisphere_grid(n_theta=256, n_phi=512):
    """Midpoint quadrature nodes over the upper hemisphere."""
    th = (np.arange(n_theta) + 0.5) * (0.5 * math.pi / n_theta)
    ph = (np.arange(n_phi) + 0.5) * (2.0 * math.pi / n_phi)
    T, P = np.meshgrid(th, ph, indexing="ij")
    w = np.sin(T) * (0.5 * math.pi / n_theta) * (2.0 * math.pi / n_phi)
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], axis=-1)
    return d.reshape(-1, 3), w.reshape(-1)


def _spec_material(model, alpha, fresnel=(1.0, 0.0)):
    p = 2.0 / (alpha * alpha) - 2.0
    return oracle.material(model, dict(diffuse=(0, 0, 0), specular=(1, 1, 1), specular_power=(p, p), fresnel=fresnel))


def _integrate(mat, wo, n_theta=192, n_phi=384):
    dirs, w = _hemisphere_grid(n_theta, n_phi)
    fsum = 0.0
    psum = 0.0
    for d, wt in zip(dirs, w):
        f, pdf = U.bsdf_eval(mat, (0, 0, 1), wo, d)
        fsum += float(f[0]) * wt
        psum += pdf * wt
    return fsum, psum


@pytest.mark.parametrize("model,alpha", [(11, 0.2), (11, 0.5), (28, 0.3), (28, 0.7)])
def test_specular_pdf_normalised_at_normal_incidence(model, alpha):
    _, psum = _integrate(_spec_material(model, alpha), (0.0, 0.0, 1.0))
    a2 = 2.0 / ((2.0 / (alpha * alpha) - 2.0) + 2.0)
    expect = 1.0 / (1.0 + a2) if model == 11 else 1.0 - math.exp(-1.0 / a2)
    assert abs(psum - expect) < 2e-3, (psum, expect)


@pytest.mark.parametrize("model,alpha,theta", [(11, 0.1, 30.0), (11, 0.4, 60.0), (28, 0.1, 30.0), (28, 0.4, 45.0)])
def test_white_lobe_energy(model, alpha, theta):
    t = math.radians(theta)
    wo = (math.sin(t), 0.0, math.cos(t))
    fsum, _ = _integrate(_spec_material(model, alpha), wo)
    assert fsum <= 1.01
    if alpha <= 0.1:
        assert fsum > 0.9, fsum


@pytest.mark.parametrize("model,alpha", [(11, 0.3), (28, 0.3), (12, 0.2), (1, 0.3), (1, 0.7)])
def test_sampler_matches_its_pdf(model, alpha):
    mat = _spec_material(model, alpha, fresnel=(0.2, 3.0) if model == 12 else (1.0, 0.0))
    mat.rec.diffuse[0] = mat.rec.diffuse[1] = mat.rec.diffuse[2] = 0.3
    t = math.radians(40.0)
    wo = (math.sin(t), 0.0, math.cos(t))
    rng = np.random.Generator(np.random.PCG64(5))
    xs = rng.uniform(0.0, 1.0, size=(20000, 3)).astype(np.float32)
    est = []
    for x in xs:
        wi, pdf = U.bsdf_sample(mat, (0, 0, 1), wo, tuple(float(v) for v in x))
        if pdf == 0.0:
            est.append(0.0)
            continue
        f, pdf2 = U.bsdf_eval(mat, (0, 0, 1), wo, wi)
        assert pdf2 == pdf
        est.append(float(f[0]) / pdf)
    est = np.array(est)
    fsum, _ = _integrate(mat, wo)
    sigma = est.std() / math.sqrt(len(est))
    assert abs(est.mean() - fsum) < 4.0 * sigma + 2e-3, (est.mean(), fsum, sigma)


@pytest.mark.parametrize("power", [1.0, 10.0, 48.0, 400.0])
def test_phong_pdf_normalised_at_normal_incidence(power):
    mat = oracle.material(1, dict(diffuse=(0, 0, 0), specular=(1, 1, 1), specular_power=(power, power)))
    _, psum = _integrate(mat, (0.0, 0.0, 1.0), n_theta=384, n_phi=64)
    expect = 1.0 - 2.0 ** (-(power + 1.0) / 2.0)
    assert abs(psum - expect) < 2e-3, (psum, expect)


@pytest.mark.parametrize("power,theta", [(10.0, 0.0), (48.0, 0.0), (48.0, 30.0), (400.0, 60.0)])
def test_phong_energy(power, theta):
    """The white lobe with F = 1 (Schlick with R0 = 1): at normal incidence
    wi makes twice the half-vector angle, so the albedo is
    (n + 8) / 4 * int_0^{pi/2} cos^n(t / 2) cos t sin t dt (1.07478 for
    n = 10, 1.03385 for n = 48); brdf-test.cc:66-85 bounds it by 1.1."""
    from scipy.integrate import quad
    t = math.radians(theta)
    mat = oracle.material(1, dict(diffuse=(0, 0, 0), specular=(1, 1, 1), specular_power=(power, power)))
    fsum, _ = _integrate(mat, (math.sin(t), 0.0, math.cos(t)), n_theta=384, n_phi=384)
    assert fsum <= 1.1, fsum
    if theta == 0.0:
        want = (power + 8.0) / 4.0 * quad(lambda x: math.cos(x / 2) ** power * math.cos(x) * math.sin(x), 0.0,
                                          0.5 * math.pi, limit=200)[0]
        assert abs(fsum - want) < 2e-3, (fsum, want)


def test_phong_half_vector_warp_matches_reference():
    """With Diffuse 0 the sampler takes the specular lobe: the half vector is
    PowerCosineSampleHemisphere(ra, rb, n) in the normal's basis (identity
    for n = +z) and wi its reflection of wo = n."""
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_phong.npz"))
    for (u, v), p, ref in zip(g["uv"], g["power"], g["power_cosine_sample_hemisphere"]):
        mat = oracle.material(1, dict(diffuse=(0, 0, 0), specular=(1, 1, 1), specular_power=(float(p), float(p))))
        wi, pdf = U.bsdf_sample(mat, (0, 0, 1), (0, 0, 1), (float(u), float(v), 0.0))
        h = ref.astype(np.float64) / np.linalg.norm(ref)
        want = 2.0 * h[2] * h - np.array([0.0, 0.0, 1.0])
        assert np.abs(wi - want).max() < 2e-5, (u, v, p, wi, want)
        # half vectors beyond 45 deg reflect below the horizon: PDF 0 there
        assert (pdf > 0.0) == (want[2] >= 0.0) or abs(want[2]) < 1e-5


@pytest.mark.parametrize("model", [11, 1])
def test_lambert_weight_is_albedo(model):
    mat = oracle.material(model, dict(diffuse=(0.7, 0.5, 0.3), specular=(0, 0, 0)))
    rng = np.random.Generator(np.random.PCG64(3))
    n = np.array([0.3, 0.4, 0.866], dtype=np.float32)
    n /= np.linalg.norm(n)
    for x in rng.uniform(0.0, 1.0, size=(200, 3)):
        wi, pdf = U.bsdf_sample(mat, n, n, tuple(float(v) for v in x))
        f, _ = U.bsdf_eval(mat, n, n, wi)
        np.testing.assert_allclose(f / pdf, [0.7, 0.5, 0.3], rtol=2e-6)


def _sggx_eval(stddev, quat, dirs):
    out = np.zeros(2, dtype=np.float32)
    res = np.zeros((len(dirs), 2))
    sd = np.asarray(stddev, dtype=np.float32)
    q = np.asarray(quat, dtype=np.float32)
    for i, d in enumerate(dirs):
        dd = np.asarray(d, dtype=np.float32)
        U.lib().orc_sggx(U.fp(sd), U.fp(q), U.fp(dd), U.fp(out))
        res[i] = out
    return res


@pytest.mark.parametrize("stddev,quat", [((0.3, 0.7, 1.0), (0, 0, 0, 1)),
                                         ((0.5, 0.25, 1.0), (0.2, -0.4, 0.1, 0.888819))])
def test_sggx_ndf_identities(stddev, quat):
    q = np.asarray(quat)
    q = q / np.linalg.norm(q)
    dirs, w = _hemisphere_grid(160, 320)
    full = np.concatenate([dirs, -dirs])
    ww = np.concatenate([w, w])
    ev = _sggx_eval(stddev, q, full)
    ndf = ev[:, 0]
    # with the reference's 0.5 factor, integral of D |w.v| over the sphere is
    # the projected area sqrt(v^T S v) (Heitz et al. 2015, eq. 5 with the
    # clamped cosine over one hemisphere)
    for v in ([0, 0, 1], [1, 0, 0], [0.6, 0.0, 0.8]):
        v = np.asarray(v, dtype=np.float64)
        proj = (ndf * np.abs(full @ v) * ww).sum()
        sigma = _sggx_eval(stddev, q, [v])[0, 1]
        assert abs(proj - sigma) < 2e-3 * max(1.0, sigma), (proj, sigma)


def test_sggx_isotropic_ndf_is_constant():
    ev = _sggx_eval((1.0, 1.0, 1.0), (0.0, 0.0, 0.0, 1.0), [[0, 0, 1], [1, 0, 0], [0.6, 0.0, 0.8], [0, -1, 0]])
    np.testing.assert_allclose(ev[:, 0], 0.5 / PI32, rtol=1e-6)
    np.testing.assert_allclose(ev[:, 1], 1.0, rtol=1e-6)


def test_sggx_surface_energy_bounded():
    mat = oracle.material(24, dict(diffuse=(0.0, 0.0, 0.0), specular=(1.0, 1.0, 1.0), specular_power=(0.4, 0.4),
                                   sggx_basis=(0.0, 0.0, 0.0, 1.0)))
    for theta in (10.0, 45.0, 75.0):
        t = math.radians(theta)
        wo = (math.sin(t), 0.0, math.cos(t))
        dirs, w = _hemisphere_grid(96, 192)
        total = 0.0
        for d, wt in zip(dirs, w):
            f, _ = U.bsdf_eval(mat, (0, 0, 1), wo, d, t=(1, 0, 0), b=(0, 1, 0), dir_density=-1.0)
            total += float(f[0]) * wt
        assert 0.0 < total <= 1.1, (theta, total)
