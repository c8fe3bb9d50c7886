"""The oracle's backbuffer resolve against the sRGB standard, the published
ACES fit and the auto-exposure definition (ray-tracing.cc:1638-1696,
2366-2375), computed here in float64 from the formulas' own statements.

Truncation to 8 bits (ToColor, spectrum.hh:792-798) is compared exactly
wherever the float64 value lies more than 1e-3 of a code from an integer
code boundary; closer values may round either way in float32.
"""
import numpy as np
import pytest

import oracle

LINEAR, SRGB, AUTO_SRGB, AUTO_HDR, ACES = range(5)


def _srgb64(v):
    """IEC 61966-2-1 encoding: 12.92 v below 0.0031308, else 1.055 v^(1/2.4) - 0.055."""
    v = np.asarray(v, dtype=np.float64)
    return np.where(v <= 0.0031308, 12.92 * v, 1.055 * np.power(np.maximum(v, 0.0031308), 1.0 / 2.4) - 0.055)


def _aces64(x):
    """Narkowicz's fit of the ACES filmic curve, clamped to [0, 1]."""
    x = np.asarray(x, dtype=np.float64)
    return np.clip(x * (2.51 * x + 0.03) / (x * (2.43 * x + 0.59) + 0.14), 0.0, 1.0)


def _codes(v):
    return 255.0 * np.clip(v, 0.0, 1.0)


def _check_channels(rgba, want):
    """rgba uint32 [n]; want float64 [n, 3] of unclamped channel values."""
    got = np.stack([(rgba >> (8 * k)) & 0xFF for k in range(3)], axis=1).astype(np.int64)
    c = _codes(want)
    safe = np.abs(c - np.round(c)) > 1e-3
    safe |= (c == 0.0) | (c == 255.0)
    assert np.all((rgba >> 24) == 255)
    assert np.array_equal(got[safe], np.floor(c[safe]).astype(np.int64))
    assert np.all(np.abs(got - np.floor(c)) <= 1)
    assert safe.mean() > 0.9


def _image(n=4096, seed=3):
    rng = np.random.default_rng(seed)
    # radiance spread over five decades, with exact zeros and values above 1
    v = np.exp(rng.uniform(np.log(1e-4), np.log(20.0), size=(n, 3)))
    v[rng.random((n, 3)) < 0.05] = 0.0
    return v


@pytest.mark.parametrize("spp", [1, 7])
def test_linear_and_srgb(spp):
    img = _image()
    acc = img * spp
    rgba, e = oracle.resolve(acc, 1.0 / spp, LINEAR)
    assert e == 0.0
    mean = (acc * (1.0 / spp)).astype(np.float32).astype(np.float64)
    _check_channels(rgba, mean)
    rgba, _ = oracle.resolve(acc, 1.0 / spp, SRGB)
    _check_channels(rgba, _srgb64(mean))


def test_known_codes():
    # 0.5 linear -> code 127; 18% grey -> sRGB 0.4614 -> code 117; 1 -> 255;
    # the sRGB break point 0.0031308 -> 0.04045 -> code 10; negatives clamp
    acc = np.array([[0.5, 1.0, -1.0], [0.18, 0.0031308, 4.0]], dtype=np.float64)
    rgba, _ = oracle.resolve(acc, 1.0, LINEAR)
    assert rgba[0] == (127 | (255 << 8) | (0 << 16) | (255 << 24))
    rgba, _ = oracle.resolve(acc, 1.0, SRGB)
    assert rgba[1] == (117 | (10 << 8) | (255 << 16) | (255 << 24))


def test_aces():
    img = _image(seed=5)
    rgba, _ = oracle.resolve(img, 1.0, ACES)
    _check_channels(rgba, _srgb64(_aces64(img.astype(np.float32))))
    # ACES(0) = 0; the curve saturates (2.51 / 2.43 > 1) and clamps to 1,
    # whose sRGB value 1.055f - 0.055f rounds to 1 - 2^-24 in float32, so
    # ToColor gives code 254 as the reference's float arithmetic does
    rgba, _ = oracle.resolve(np.array([[0.0, 1e6, 0.18]]), 1.0, ACES)
    assert rgba[0] & 0xFF == 0 and (rgba[0] >> 8) & 0xFF == 254
    assert (rgba[0] >> 16) & 0xFF == int(255 * _srgb64(_aces64(0.18)))


def _exposure64(mean):
    ok = np.all(np.isfinite(mean), axis=1)
    g = np.exp(np.log(1e-3 + mean[ok]).sum(axis=0) / mean.shape[0])
    return 0.18 / (0.2126 * g[0] + 0.7152 * g[1] + 0.0722 * g[2])


def test_auto_exposure_constant_image():
    # a grey image of value v: the log average is v + 1e-3, the luminance
    # weights sum to one, so the exposure is 0.18 / (v + 1e-3)
    for v in (0.01, 0.5, 3.0):
        acc = np.full((1000, 3), v)
        _, e = oracle.resolve(acc, 1.0, AUTO_SRGB)
        assert e == pytest.approx(0.18 / (v + 1e-3), rel=2e-6)


def test_auto_exposure_modes():
    img = _image(seed=9)
    mean = img.astype(np.float32).astype(np.float64)
    want_e = _exposure64(mean)
    rgba, e = oracle.resolve(img, 1.0, AUTO_SRGB)
    assert e == pytest.approx(want_e, rel=1e-5)
    _check_channels(rgba, _srgb64(np.float32(e) * mean))
    rgba, e2 = oracle.resolve(img, 1.0, AUTO_HDR)
    assert e2 == e
    x = np.float32(e) * mean
    _check_channels(rgba, _srgb64(x / (1.0 + x)))


def test_auto_exposure_skips_non_finite_pixels():
    # the log sum takes finite pixels only, the mean divides by every pixel
    img = _image(n=512, seed=11)
    img[7] = [np.inf, 0.5, 0.5]
    img[300] = [0.2, np.nan, 0.1]
    _, e = oracle.resolve(img, 1.0, AUTO_SRGB)
    assert e == pytest.approx(_exposure64(img.astype(np.float32).astype(np.float64)), rel=1e-5)


def test_exposure_override():
    img = _image(seed=13)
    a, e = oracle.resolve(img, 1.0, AUTO_HDR)
    b, e_same = oracle.resolve(img, 1.0, AUTO_HDR, exposure=e)
    assert np.array_equal(a, b) and e_same == e
    c, _ = oracle.resolve(img, 1.0, AUTO_HDR, exposure=2.0 * e)
    assert (c != a).any()
