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
