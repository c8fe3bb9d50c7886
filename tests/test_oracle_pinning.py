"""Pins the CPU oracle before it is trusted as the checker.

- Philox-4x32-10 against the known-answer vectors of Salmon et al. (SC'11)
  as published with Random123 (counter/key all zero, all ones, and the
  pi-digit vector).
- The shared elementary functions (tr_fmath.h) against float64 libm.
- The light/hemisphere warps, Fresnel terms and basis construction
  against golden vectors produced by the reference's own code
  (tests/golden/ref_sampling.npz, from oracle/_ref; see make_golden.py).
  The reference computes sin/cos with a tan-half-angle formula
  (functions.hh:322-329) while the oracle uses tr_sincos_turns, so warps
  that go through an angle are compared with a 1e-5 tolerance; those that
  do not are compared to 1-2 ulp.
"""
import os

import numpy as np
import pytest

import _oracle_units as U
import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_sampling.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLDEN)


def test_philox_known_answers():
    assert U.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert U.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert U.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_uniform_is_open_interval():
    lb = U.lib()
    for x in (0, 1, 255, 256, 0xFFFFFFFF, 0xFFFFFF00, 0x80000000):
        u = lb.orc_u01f(x)
        assert 0.0 < u < 1.0
    assert lb.orc_u01f(0) == 2.0 ** -24
    assert lb.orc_u01f(0xFFFFFFFF) == 1.0 - 2.0 ** -24


def _ulp_err(got, want):
    want32 = want.astype(np.float32)
    ulp = np.abs(np.nextafter(want32, np.float32(np.inf)) - want32).astype(np.float64)
    return np.abs(got.astype(np.float64) - want) / ulp


def test_sincos_turns_accuracy():
    t = np.linspace(0.0, 1.0, 200001, dtype=np.float32)
    s = U.fmath(0, t)
    c = U.fmath(1, t)
    a = 2.0 * np.pi * t.astype(np.float64)
    ms = np.abs(np.sin(a)) > 1e-4
    mc = np.abs(np.cos(a)) > 1e-4
    assert _ulp_err(s, np.sin(a))[ms].max() <= 2.0
    assert _ulp_err(c, np.cos(a))[mc].max() <= 2.0
    assert np.abs(s - np.sin(a)).max() < 2e-7 and np.abs(c - np.cos(a)).max() < 2e-7
    # quadrant points are exact
    q = U.fmath(0, np.array([0.0, 0.25, 0.5, 0.75], dtype=np.float32))
    assert list(q) == [0.0, 1.0, 0.0, -1.0]


def test_log_exp_accuracy():
    x = np.exp(np.linspace(-87.0, 88.0, 100001)).astype(np.float32)
    assert _ulp_err(U.fmath(2, x), np.log(x.astype(np.float64)))[np.abs(np.log(x)) > 1e-3].max() <= 2.0
    x = np.geomspace(1e-44, 1e-38, 1001).astype(np.float32)  # denormals
    assert np.abs(U.fmath(2, x) - np.log(x.astype(np.float64))).max() < 1e-5
    y = np.linspace(-87.0, 88.0, 100001).astype(np.float32)
    assert _ulp_err(U.fmath(3, y), np.exp(y.astype(np.float64))).max() <= 2.0
    assert U.fmath(3, np.array([-200.0, 100.0], dtype=np.float32)).tolist() == [0.0, np.inf]
    assert U.fmath(2, np.array([0.0, 1.0], dtype=np.float32)).tolist() == [-np.inf, 0.0]


def test_cosine_and_uniform_hemisphere_match_reference(gold):
    """GGX with Specular 0 samples CosineSampleHemisphere, the SGGX surface
    model UniformSampleHemisphere (sampling3.hh:55-76); for n = +z the basis
    is the identity (matrix3.hh:99-107)."""
    lam = oracle.material(11, dict(diffuse=(0.5, 0.5, 0.5), specular=(0, 0, 0)))
    sg = oracle.material(24, dict(diffuse=(0.5, 0.5, 0.5), specular=(0, 0, 0), specular_power=(0.3, 0.3)))
    uv = gold["warp_uv"]
    for i in range(8, len(uv)):  # exact 0/1 corners are outside the sampler's (0, 1) inputs
        u, v = float(uv[i, 0]), float(uv[i, 1])
        wi, _ = U.bsdf_sample(lam, (0, 0, 1), (0, 0, 1), (u, v, 0.5))
        ref = gold["cosine_sample_hemisphere"][i].astype(np.float64)
        ref /= np.linalg.norm(ref)
        assert np.abs(wi - ref).max() < 1e-5, (i, wi, ref)
        wi, pdf = U.bsdf_sample(sg, (0, 0, 1), (0, 0, 1), (u, v, 0.5))
        ref = gold["uniform_sample_hemisphere"][i].astype(np.float64)
        ref /= np.linalg.norm(ref)
        assert np.abs(wi - ref).max() < 1e-5
        assert pdf == np.float32(1.0 / (2.0 * np.float32(3.141592)))


def test_sphere_light_matches_reference(gold):
    lb = U.lib()
    n = len(gold["sphere_radius"])
    out = np.zeros(5, dtype=np.float32)
    for i in range(n):
        u, v = gold["warp_uv"][i]
        if u <= 0.0 or u >= 1.0 or v <= 0.0 or v >= 1.0:
            continue
        pos, c, r = gold["sphere_pos"][i], gold["sphere_center"][i], float(gold["sphere_radius"][i])
        lb.orc_sample_sphere_light(U.fp(pos), U.fp(c), r, float(u), float(v), U.fp(out))
        assert np.abs(out[:3] - gold["sphere_dir"][i]).max() < 2e-5
        assert abs(out[3] - gold["sphere_dist"][i]) <= 2e-5 * max(1.0, abs(gold["sphere_dist"][i]))
        assert out[4] == gold["sphere_pdf"][i]
        d = gold["sphere_dir"][i].copy()
        assert lb.orc_sphere_light_pdf(U.fp(pos), U.fp(d), U.fp(c), r) == gold["sphere_pdf_eval"][i]


def test_triangle_light_matches_reference(gold):
    lb = U.lib()
    out = np.zeros(5, dtype=np.float32)
    for i in range(len(gold["tri_pos"])):
        u, v = gold["warp_uv"][i]
        tv, pos = gold["tri_verts"][i], gold["tri_pos"][i]
        lb.orc_sample_triangle_area(U.fp(pos), U.fp(tv[0].copy()), U.fp(tv[1].copy()), U.fp(tv[2].copy()),
                                    float(u), float(v), U.fp(out))
        np.testing.assert_array_equal(out[:3], gold["tri_dir"][i])
        assert out[3] == gold["tri_dist"][i] and out[4] == gold["tri_pdf"][i]
        a = lb.orc_projected_triangle_area(U.fp(pos), U.fp(tv[0].copy()), U.fp(tv[1].copy()), U.fp(tv[2].copy()))
        assert a == gold["tri_proj_area"][i]


def test_fresnel_matches_reference(gold):
    lb = U.lib()
    for e, k, c, fc, fs in zip(gold["fresnel_eta"], gold["fresnel_k"], gold["fresnel_cos"],
                               gold["fresnel_conductor"], gold["fresnel_schlick"]):
        assert lb.orc_fresnel_conductor(float(e), float(k), float(c)) == fc
        # reference: pow(1 - cos, 5.0f); oracle: x^2 x^2 x
        s = lb.orc_fresnel_schlick(float(np.float32(float(e) / 3.0)), float(c))
        assert abs(s - fs) <= 4e-7 * max(1.0, abs(fs))
    for e, c, ct, fd in zip(gold["dielectric_eta"], gold["fresnel_cos"], gold["cos_transmittance"],
                            gold["fresnel_dielectric"]):
        assert lb.orc_fresnel_dielectric(float(e), float(c), float(ct)) == fd


def test_basis_matches_reference(gold):
    lb = U.lib()
    out = np.zeros(9, dtype=np.float32)
    for n, b in zip(gold["basis_normal"], gold["basis"]):
        lb.orc_make_basis(U.fp(n.copy()), U.fp(out))
        np.testing.assert_array_equal(out, b)
