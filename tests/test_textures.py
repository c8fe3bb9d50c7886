"""Texture maps and the global cube map in the oracle (oracle/orc_texture.h),
pinned by what the sampling rules fix (graphics/texture.hh:470-520,
src/graphics/cube-map.cc:73-110, graphics/cube-map.hh:35-77):

- nearest and bilinear lookups at texel centres return the texels exactly,
  bilinear half-way between two texels their mean, coordinates one period
  apart the same value (wrapping), and a linear ramp interpolates linearly;
- 8-bit spectra decode with the IEC 61966-2-1 curve (float64 here), RG
  and RGBA channels as byte / 255;
- a cube-map lookup along each axis reads the centre of that axis' face
  (+x, -x, +y, -y, +z, -z), faces stored with their rows reversed;
- a Blinn-Phong surface with a constant DiffuseMap renders exactly as the
  same surface with that constant as Diffuse, and a cube map of one
  colour exactly as a background of that colour.
"""
import numpy as np
import pytest

import oracle
from tempestrenderer_amd import workloads

SPECTRUM, RG, RGBA = 0, 1, 2
RG8, RGBA8, RGBA32F = 1, 2, 6


def _scene():
    return oracle.OracleScene(4, 4, np.eye(4, dtype=np.float32).reshape(16))


def _srgb_decode64(b):
    c = np.asarray(b, dtype=np.float64) / 255.0
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def test_nearest_and_centres():
    s = _scene()
    rng = np.random.default_rng(1)
    tex = rng.uniform(0, 1, size=(8, 16, 4)).astype(np.float32)
    hn = s.bindTexture(tex, RGBA32F, 1)
    hb = s.bindTexture(tex, RGBA32F, 0)
    for y in range(8):
        for x in range(16):
            u, v = (x + 0.5) / 16, (y + 0.5) / 8
            assert np.array_equal(s.textureSample(hn, u, v, RGBA), tex[y, x])
            assert np.array_equal(s.textureSample(hb, u, v, RGBA), tex[y, x])
            # one period further: the same texel
            assert np.array_equal(s.textureSample(hb, u + 1.0, v - 2.0, RGBA), tex[y, x])


def test_bilinear_midpoints_and_ramp():
    s = _scene()
    tex = np.zeros((4, 8, 4), dtype=np.float32)
    tex[..., 0] = np.arange(8)[None, :]
    tex[..., 1] = np.arange(4)[:, None] * 10.0
    h = s.bindTexture(tex, RGBA32F, 0)
    # half-way between texels 2 and 3 of a row: the mean
    got = s.textureSample(h, 3.0 / 8, 1.5 / 4, RGBA)
    assert got[0] == np.float32(2.5) and got[1] == np.float32(10.0)
    # inside the ramp (away from the wrap seam) bilinear is exact
    for u in np.linspace(0.1, 0.85, 13):
        got = s.textureSample(h, u, 0.5, RGBA)
        assert abs(got[0] - (8 * u - 0.5)) < 1e-5
    # across the seam the wrap blends the last and first texels
    got = s.textureSample(h, 0.0, 0.5, RGBA)
    assert abs(got[0] - 3.5) < 1e-6  # half of texel 7 and half of texel 0


def test_channel_decoding():
    s = _scene()
    b = np.arange(256, dtype=np.uint32)
    rgba = (b | ((255 - b) << 8) | ((b // 2) << 16) | (7 << 24)).reshape(16, 16)
    h = s.bindTexture(rgba, RGBA8, 1)
    rg = np.stack([b, 255 - b], -1).astype(np.uint8).reshape(16, 16, 2)
    h2 = s.bindTexture(rg, RG8, 1)
    for i in range(256):
        u, v = (i % 16 + 0.5) / 16, (i // 16 + 0.5) / 16
        sp = s.textureSample(h, u, v, SPECTRUM)
        want = _srgb_decode64([i, 255 - i, i // 2])
        assert np.abs(sp[:3] - want).max() < 1e-6 * max(1.0, want.max()) + 1e-9  # float32 evaluation
        assert s.textureSample(h, u, v, RGBA).tolist() == [np.float32(i / 255.0), np.float32((255 - i) / 255.0),
                                                          np.float32((i // 2) / 255.0), np.float32(7 / 255.0)]
        assert s.textureSample(h2, u, v, RG)[:2].tolist() == [np.float32(i / 255.0), np.float32((255 - i) / 255.0)]


def _faces(side=24):
    faces = []
    for f in range(6):
        img = np.zeros((side, side), dtype=np.uint32)
        img[:] = (40 * f + 10) | ((200 - 30 * f) << 8) | ((17 * f) << 16) | (255 << 24)
        faces.append(img)
    return faces


def test_cube_map_axes_pick_faces():
    s = _scene()
    faces = _faces()
    s.setGlobalCubeMap(faces)
    axes = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    for f, d in enumerate(axes):
        want = _srgb_decode64([40 * f + 10, 200 - 30 * f, 17 * f])
        assert np.abs(s.cubeMapSample(d) - want).max() < 1e-6


def test_cube_map_rows_are_flipped():
    """Face +z: t = -y, so a direction tilted towards +y reads a low stored
    row, which holds a high row of the face image as given."""
    s = _scene()
    faces = _faces()
    top = np.full((24, 24), 0xFF0000FF, dtype=np.uint32)   # red everywhere
    top[12:] = 0xFFFF0000                                   # image rows 12..23 blue
    faces[4] = top
    s.setGlobalCubeMap(faces)
    up = s.cubeMapSample((0.0, 0.5, 1.0))
    down = s.cubeMapSample((0.0, -0.5, 1.0))
    assert up[2] > 0.9 and up[0] < 0.1      # blue: the image's lower rows
    assert down[0] > 0.9 and down[2] < 0.1  # red


def _plate(diffuse, diffuse_map=None, cube=None, background=(0.3, 0.3, 0.3)):
    vpi = workloads.perspective_view_proj_inv((0.0, 1.5, 1.5), (0, 0, 0), (0, 1, 0), 50.0, 1.0)
    s = oracle.OracleScene(16, 16, vpi)
    fields = dict(diffuse=diffuse, specular=(0.1, 0.1, 0.1), specular_power=(12.0, 12.0))
    if diffuse_map is not None:
        # nearest: bilinear weights of a constant texture round within an ulp
        fields["diffuse_map"] = s.bindTexture(diffuse_map, RGBA8, 1)
    if cube is not None:
        s.setGlobalCubeMap(cube)
    qv, qt = workloads._textured_quad((0, 0, 0), (0, 0, 1), (1, 0, 0), tc_scale=2.0)
    from tempestrenderer_amd.scene import Submesh
    m = oracle.material(1, fields)
    s.addTriangleMesh(workloads.IDENTITY, [Submesh(m, 6, 0, 0, 32)], qt, qv)
    s.addLightSource(s.make_light(kind=0, direction=(0.3, 0.9, 0.3), radiance=(1.0, 1.0, 1.0)))
    s.commitScene()
    s.setSamplesCamera(2)
    s.setMaxRayDepth(3)
    s.setBackgroundSpectrum(background)
    return s.renderPaths(np.arange(256, dtype=np.uint32))[0]


def test_constant_map_equals_constant_diffuse():
    b = np.array([150, 90, 30])
    lin = _srgb_decode64(b).astype(np.float32)
    texel = np.uint32(b[0] | (b[1] << 8) | (b[2] << 16) | (255 << 24))
    with_map = _plate((0.9, 0.9, 0.9), diffuse_map=np.full((4, 4), texel, dtype=np.uint32))
    s = oracle.OracleScene(4, 4, np.eye(4, dtype=np.float32).reshape(16))
    h = s.bindTexture(np.full((4, 4), texel, dtype=np.uint32), RGBA8, 0)
    decoded = s.textureSample(h, 0.3, 0.7, SPECTRUM)[:3]
    assert np.abs(decoded - lin).max() < 1e-6
    plain = _plate(tuple(float(x) for x in decoded))
    assert np.array_equal(with_map, plain)


def test_uniform_cube_map_equals_background():
    faces = [np.full((20, 20), 0xFF4080C0, dtype=np.uint32)] * 6
    s = oracle.OracleScene(4, 4, np.eye(4, dtype=np.float32).reshape(16))
    s.setGlobalCubeMap(faces)
    c = s.cubeMapSample((0.3, 0.4, 0.5))
    assert np.abs(c - _srgb_decode64([0xC0, 0x80, 0x40])).max() < 1e-6
    a = _plate((0.5, 0.5, 0.5), cube=faces)
    b = _plate((0.5, 0.5, 0.5), background=tuple(float(x) for x in c))
    # the clamped bilinear weights of a constant face round within an ulp
    np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-8)
    assert (a != 0).all()


def test_textured_workload_renders():
    d = workloads.textured(16, 16, 2, 3)
    o = oracle.OracleScene(d.width, d.height, d.vpi)
    d.populate(o, oracle.material)
    o.commitScene()
    rgb, _ = o.renderPaths(np.arange(d.pixels, dtype=np.uint32))
    assert np.isfinite(rgb).all() and rgb.mean() > 0.05
