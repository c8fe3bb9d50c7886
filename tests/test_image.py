"""Image files of the backbuffer through the C ABI (host code, no GPU).

TGA is pinned byte for byte by the TGA 2.0 layout the reference writer
fills in (src/image/tga-image.cc:247-309: 18-byte header, type 2, 32 bpp,
descriptor 0, B G R A pixels, rows from the bottom) and by an independent
decoder (Pillow); PNG by Pillow decoding the same pixels the reference's
stb_image_write call stores (src/image/png-image.cc:264-266: last row
first).
"""
import io
import struct

import numpy as np
import pytest

from tempestrenderer_amd import _abi
from tempestrenderer_amd._abi import TempestError
from tempestrenderer_amd.scene import encode_image, save_image

Image = pytest.importorskip("PIL.Image")


def _picture(h=37, w=53, seed=2, alpha=None):
    rng = np.random.default_rng(seed)
    px = rng.integers(0, 1 << 32, size=(h, w), dtype=np.uint64).astype(np.uint32)
    if alpha is not None:
        px = (px & np.uint32(0x00FFFFFF)) | np.uint32(alpha << 24)
    return px


def _channels(px):
    """[h, w] packed RGBA (R low byte) -> [h, w, 4] uint8."""
    return np.stack([(px >> (8 * k)) & 0xFF for k in range(4)], axis=-1).astype(np.uint8)


def test_tga_layout():
    px = _picture()
    data = encode_image(px, _abi.TR_IMAGE_TGA)
    h, w = px.shape
    assert len(data) == 18 + 4 * w * h
    hdr = struct.unpack("<BBBHHBHHHHBB", data[:18])
    assert hdr == (0, 0, 2, 0, 0, 0, 0, 0, w, h, 32, 0)
    body = np.frombuffer(data[18:], dtype=np.uint8).reshape(h, w, 4)
    c = _channels(px)
    assert np.array_equal(body[..., 0], c[..., 2]) and np.array_equal(body[..., 2], c[..., 0])
    assert np.array_equal(body[..., 1], c[..., 1]) and np.array_equal(body[..., 3], c[..., 3])


def test_tga_decodes_bottom_up():
    px = _picture(alpha=255)
    img = Image.open(io.BytesIO(encode_image(px, _abi.TR_IMAGE_TGA)))
    got = np.asarray(img.convert("RGBA"))
    assert np.array_equal(got, _channels(px)[::-1])


def test_png_decodes_same_pixels():
    px = _picture(seed=4)
    img = Image.open(io.BytesIO(encode_image(px, _abi.TR_IMAGE_PNG)))
    assert img.mode == "RGBA" and img.size == (px.shape[1], px.shape[0])
    assert np.array_equal(np.asarray(img), _channels(px)[::-1])


def test_png_and_tga_agree():
    px = _picture(h=64, w=64, seed=6, alpha=255)
    a = np.asarray(Image.open(io.BytesIO(encode_image(px, _abi.TR_IMAGE_TGA))).convert("RGBA"))
    b = np.asarray(Image.open(io.BytesIO(encode_image(px, _abi.TR_IMAGE_PNG))))
    assert np.array_equal(a, b)


def test_save_by_extension(tmp_path):
    px = _picture(h=8, w=5)
    save_image(tmp_path / "a.tga", px)
    assert (tmp_path / "a.tga").read_bytes() == encode_image(px, _abi.TR_IMAGE_TGA)
    save_image(tmp_path / "b.PNG", px)  # the extension is lower-cased first
    assert (tmp_path / "b.PNG").read_bytes() == encode_image(px, _abi.TR_IMAGE_PNG)
    for name in ("c.exr", "d.bmp", "noext", "dir.tga/e"):
        with pytest.raises(TempestError):
            save_image(tmp_path / name, px)
    assert not (tmp_path / "c.exr").exists()


def test_bad_arguments():
    with pytest.raises(TempestError):
        encode_image(np.zeros((0, 4), dtype=np.uint32))
    with pytest.raises(TempestError):
        encode_image(np.zeros((1, 70000), dtype=np.uint32), _abi.TR_IMAGE_TGA)
    with pytest.raises(TempestError):
        encode_image(np.zeros((2, 2), dtype=np.uint32), 7)


# ---------------- readers ----------------

from tempestrenderer_amd.scene import load_image  # noqa: E402

GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")


def _rgba_bytes(a):
    return a.view(np.uint8).reshape(a.shape[0], a.shape[1], 4)


@pytest.mark.parametrize("name", ["cube/default.png", "storforsen4/posx-256.png", "storforsen4/negy-256.png"])
def test_png_fixtures_match_pillow(name):
    path = GOLDEN + "/" + name
    a, fmt = load_image(path)
    assert fmt == _abi.TR_TEX_RGBA8_UNORM
    assert np.array_equal(_rgba_bytes(a), np.asarray(Image.open(path).convert("RGBA")))


def _pil_png(tmp_path, img, name, **kw):
    p = tmp_path / name
    img.save(p, **kw)
    return str(p)


def test_png_colour_types_match_pillow(tmp_path):
    """RGB (alpha 255), grey-alpha (grey in R, G, B), 8-bit palettes with and
    without tRNS, a 4-bit palette and 1-bit grey, each against Pillow; every
    row filter occurs in Pillow's adaptive filtering of these images."""
    rng = np.random.default_rng(3)
    rgb = rng.integers(0, 256, size=(37, 29, 3), dtype=np.uint8)
    rgb[:, :10] = rgb[:1, :10]  # runs, so that filters differ per row
    p = _pil_png(tmp_path, Image.fromarray(rgb, "RGB"), "rgb.png")
    a, fmt = load_image(p)
    want = np.concatenate([rgb, np.full((37, 29, 1), 255, np.uint8)], -1)
    assert fmt == _abi.TR_TEX_RGBA8_UNORM and np.array_equal(_rgba_bytes(a), want)
    la = rng.integers(0, 256, size=(20, 33, 2), dtype=np.uint8)
    p = _pil_png(tmp_path, Image.fromarray(la, "LA"), "la.png")
    a, _ = load_image(p)
    assert np.array_equal(_rgba_bytes(a), np.asarray(Image.open(p).convert("RGBA")))
    pal = Image.fromarray(rgb, "RGB").quantize(colors=200)
    for name, kw in (("p8.png", {}), ("p8t.png", {"transparency": 5})):
        p = _pil_png(tmp_path, pal, name, **kw)
        a, _ = load_image(p)
        assert np.array_equal(_rgba_bytes(a), np.asarray(Image.open(p).convert("RGBA")))
    pal4 = Image.fromarray(rgb, "RGB").quantize(colors=12)
    p = _pil_png(tmp_path, pal4, "p4.png", bits=4)
    a, _ = load_image(p)
    assert np.array_equal(_rgba_bytes(a), np.asarray(Image.open(p).convert("RGBA")))


def test_png_grey_is_r8(tmp_path):
    rng = np.random.default_rng(4)
    g = rng.integers(0, 256, size=(17, 23), dtype=np.uint8)
    a, fmt = load_image(_pil_png(tmp_path, Image.fromarray(g, "L"), "g.png"))
    assert fmt == _abi.TR_TEX_R8_UNORM and np.array_equal(a, g)
    bits = rng.integers(0, 2, size=(9, 13)).astype(bool)
    a, fmt = load_image(_pil_png(tmp_path, Image.fromarray(bits), "b.png"))
    assert fmt == _abi.TR_TEX_R8_UNORM and np.array_equal(a, bits.astype(np.uint8) * 255)


def test_png_rejected_kinds(tmp_path):
    img = Image.fromarray(np.zeros((8, 8, 3), np.uint8), "RGB")
    with pytest.raises(TempestError):
        load_image(str(tmp_path / "none.png"))
    p16 = _pil_png(tmp_path, Image.fromarray(np.zeros((8, 8), np.uint16)), "w.png")
    with pytest.raises(TempestError):
        load_image(p16)
    bad = tmp_path / "bad.png"
    data = bytearray(open(_pil_png(tmp_path, img, "ok.png"), "rb").read())
    data[40] ^= 0xFF  # inside IDAT: CRC mismatch
    bad.write_bytes(bytes(data))
    with pytest.raises(TempestError):
        load_image(str(bad))


def test_tga_round_trip_and_pillow(tmp_path):
    px = _picture(h=21, w=34, seed=8)
    save_image(tmp_path / "w.tga", px)
    a, fmt = load_image(str(tmp_path / "w.tga"))
    assert fmt == _abi.TR_TEX_RGBA8_UNORM and np.array_equal(a, px)  # bottom-up as written
    rgb = np.random.default_rng(2).integers(0, 256, size=(12, 19, 3), dtype=np.uint8)
    rgb[3:9] = rgb[3]  # runs for the run-length coder
    for name, kw in (("p.tga", {}), ("r.tga", {"compression": "tga_rle"})):
        Image.fromarray(rgb, "RGB").save(tmp_path / name, **kw)
        a, _ = load_image(str(tmp_path / name))
        got = _rgba_bytes(a)
        # texture rows run bottom-up whatever the file's origin (a top-down
        # file is flipped); Pillow's array is top-down
        assert np.array_equal(got[..., :3], rgb[::-1]) and (got[..., 3] == 255).all()


def test_obj_loads_map_kd():
    from tempestrenderer_amd.scene import LoadObjFileStaticRTGeometry
    b = LoadObjFileStaticRTGeometry(GOLDEN + "/cube/cube.obj")
    assert b.IgnoredMaps == 0 and len(b.Textures) == 1
    tex, fmt = b.Textures[0]
    assert fmt == _abi.TR_TEX_RGBA8_UNORM and tex.shape == (128, 128)
    assert all(m.rec.diffuse_map == 1 for m in b.Materials)
    assert np.array_equal(_rgba_bytes(tex), np.asarray(Image.open(GOLDEN + "/cube/default.png").convert("RGBA")))
