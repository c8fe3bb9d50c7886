"""The SGGX pseudo-volume surface of the oracle (model 25:
SGGXMicroFlakePseudoVolumeBRDF, SGGXMicroFlakeSurfacePDF and
SGGXMicroFlakeSurfaceSampleIncidentLight, ggx-models-impl.hh:621-842),
pinned by:

- its internal random numbers: IntelFastRand is the LCG of the Microsoft C
  runtime's rand() (x <- 214013 x + 2531011, bits 16..30), whose sequence
  from seed 1 is the published 41, 18467, 6334, 26500, ...;
- Depth 0 (no walks) leaves the single-scattering term, which is the SGGX
  surface model (24, already pinned) with Fresnel 1 and a unit quaternion;
- a specular channel at or under 1e-3 stops every walk before it starts
  (the reference's `throughput > 1e-3f` holds only when every channel
  passes), so any Depth gives the Depth-0 value;
- the walks add non-negative radiance, and at Depth > 0 some;
- its PDF is twice the micro-flake PDF (model 22) in the same frame with
  the identity quaternion, and the sampler returns the PDF function's value
  at the direction it draws (0 below the surface).
"""
import numpy as np
import pytest

import _oracle_units as U
import oracle

N = (0.0, 0.0, 1.0)
T = (1.0, 0.0, 0.0)
B = (0.0, 1.0, 0.0)


def _pv(depth, count=64, specular=(0.8, 0.7, 0.6), stddev=(0.3, 0.5), quat=(0.0, 0.0, 0.0, 1.0)):
    return oracle.material(25, dict(diffuse=(0.1, 0.2, 0.1), specular=specular, specular_power=stddev,
                                    sggx_basis=quat, sggx_depth=depth, sggx_sample_count=count))


def _dirs(seed, n):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d[:, 2] = np.abs(d[:, 2]) + 0.1
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def test_lcg_is_the_msvc_rand_sequence():
    out = np.zeros(10, dtype=np.float32)
    U.lib().orc_lcg15(1, 10, U.fp(out))
    want = np.array([41, 18467, 6334, 26500, 19169, 15724, 11478, 29358, 26962, 24464], dtype=np.float32) / 32767.0
    np.testing.assert_array_equal(out, want.astype(np.float32))


def test_depth_zero_is_the_sggx_surface_single_scattering():
    q = np.array([0.2, -0.4, 0.1, 0.8])
    q = tuple(float(v) for v in q / np.linalg.norm(q))
    pv = _pv(0, quat=q)
    surf = oracle.material(24, dict(diffuse=(0.1, 0.2, 0.1), specular=(0.8, 0.7, 0.6), specular_power=(0.3, 0.5),
                                    sggx_basis=q, fresnel=(1.0, 0.0)))
    for wo, wi in zip(_dirs(1, 64), _dirs(2, 64)):
        fa, _ = U.bsdf_eval(pv, N, wo, wi, t=T, b=B, dir_density=0.7)
        fb, _ = U.bsdf_eval(surf, N, wo, wi, t=T, b=B, dir_density=0.7)
        np.testing.assert_allclose(fa, fb, rtol=2e-6, atol=1e-9)


def test_dark_channel_stops_the_walks():
    for wo, wi in zip(_dirs(3, 32), _dirs(4, 32)):
        f0, _ = U.bsdf_eval(_pv(0, specular=(0.8, 0.001, 0.6)), N, wo, wi, t=T, b=B, dir_density=0.7)
        f3, _ = U.bsdf_eval(_pv(3, specular=(0.8, 0.001, 0.6)), N, wo, wi, t=T, b=B, dir_density=0.7)
        np.testing.assert_array_equal(f0, f3)


def test_walks_add_radiance():
    gained = 0
    for wo, wi in zip(_dirs(5, 48), _dirs(6, 48)):
        f0, _ = U.bsdf_eval(_pv(0), N, wo, wi, t=T, b=B)
        f2, _ = U.bsdf_eval(_pv(2), N, wo, wi, t=T, b=B)
        assert np.all(f2 >= f0 * (1.0 - 1e-6))
        gained += bool(np.any(f2 > f0 * 1.001))
    assert gained > 24


def test_pdf_is_twice_the_flake_pdf_and_matches_the_sampler():
    flake = oracle.material(22, dict(sggx_stddev=(0.3, 0.5, 1.0)))
    pv = _pv(2)
    for wo, wi in zip(_dirs(7, 64), _dirs(8, 64)):
        _, p25 = U.bsdf_eval(pv, N, wo, wi, t=T, b=B, dir_density=0.6)
        _, p22 = U.bsdf_eval(flake, N, wo, wi, t=T, b=B, dir_density=0.6)
        assert p25 == pytest.approx(2.0 * p22, rel=1e-6)
    rng = np.random.default_rng(9)
    hits = 0
    for wo in _dirs(10, 200):
        xi = tuple(float(v) for v in rng.uniform(size=3))
        wi, pdf = U.bsdf_sample(pv, N, wo, xi, t=T, b=B, dir_density=0.6)
        if pdf == 0.0:
            assert wi[2] <= 0.0
            continue
        hits += 1
        assert wi[2] > 0.0
        _, p2 = U.bsdf_eval(pv, N, wo, wi, t=T, b=B, dir_density=0.6)
        assert pdf == pytest.approx(p2, rel=1e-4)
    # the reference's flake sampler maps S into the frame of wo as B S B^T
    # (not B^T S B), so its directions do not follow its PDF and only part of
    # them lands above the surface; what it returns is still the PDF's value
    assert hits > 20
