"""Image files on either side of the path (rd_image_load / rd_image_save_*),
against TextureUtility's behaviour (Radiantium.Core/TextureUtility.cs:54-203)
through Magick.NET-Q16-HDRI.

Inputs come from tests/golden/image_fixtures.py, whose OpenEXR / Radiance /
PFM encoders are written from the file specifications independently of the
C++ decoders (no OpenEXR library is installed and the reference's .exr
assets are absent).  PNG output is read back with Pillow, an independent
decoder.  Host code only: no GPU needed."""
import os
import struct
import zlib
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import image_fixtures as fx  # noqa: E402

from radiantium_amd import abi  # noqa: E402
from radiantium_amd.native import RadiantiumError, image_load, image_save_exr, image_save_pfm, image_save_png  # noqa: E402


def as_rows(cb):
    """ColorBuffer [x, y, c] loaded with flip_y=False -> rows top first [y, x, c]."""
    return np.transpose(cb, (1, 0, 2))


EXR_CASES = [
    ("none_half", fx.NONE, fx.HALF, "RGB", (5, 7)),
    ("rle_half_rgba", fx.RLE, fx.HALF, "RGBA", (9, 11)),
    ("zips_float", fx.ZIPS, fx.FLOAT, "RGB", (6, 13)),
    ("zip_half", fx.ZIP, fx.HALF, "RGB", (37, 21)),       # 16-line blocks, a partial last block
    ("zip_uint", fx.ZIP, fx.UINT, "RGB", (17, 9)),
    ("piz_half", fx.PIZ, fx.HALF, "RGB", (45, 13)),       # 32-line blocks, partial, odd sizes
    ("piz_half_rgba", fx.PIZ, fx.HALF, "RGBA", (32, 32)),
    ("piz_float", fx.PIZ, fx.FLOAT, "RGB", (19, 23)),     # two 16-bit words per sample
    ("piz_uint_wide", fx.PIZ, fx.UINT, "RGB", (33, 40)),  # many distinct values: the 16-bit wavelet
    ("piz_y", fx.PIZ, fx.HALF, "Y", (8, 8)),
    ("none_y", fx.NONE, fx.FLOAT, "Y", (3, 4)),
]


@pytest.mark.parametrize("name,comp,ptype,chans,shape", EXR_CASES, ids=[c[0] for c in EXR_CASES])
def test_exr_decode(tmp_path, name, comp, ptype, chans, shape):
    h, w = shape
    ch = fx.sample_channels(h, w, zlib.crc32(name.encode()) % 1000, ptype, chans)
    p = tmp_path / f"{name}.exr"
    fx.write_exr(str(p), ch, comp)
    got = as_rows(image_load(p, flip_y=False, cast_to_linear=False))
    want = fx.exr_expected(ch)
    assert got.shape == want.shape
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_exr_piz_zero_image_and_runs(tmp_path):
    # all zeros: no bitmap bytes (min > max) and one long Huffman run
    ch = {n: (fx.HALF, np.zeros((40, 50), np.float32)) for n in "RGB"}
    fx.write_exr(str(tmp_path / "z.exr"), ch, fx.PIZ)
    got = as_rows(image_load(tmp_path / "z.exr", flip_y=False, cast_to_linear=False))
    np.testing.assert_array_equal(got, fx.exr_expected(ch))
    # long runs of one value, then noise
    a = np.full((33, 70), 0.25, np.float32)
    a[20:] = np.random.default_rng(5).random((13, 70)).astype(np.float32)
    ch = {n: (fx.HALF, a * (k + 1)) for k, n in enumerate("RGB")}
    fx.write_exr(str(tmp_path / "r.exr"), ch, fx.PIZ)
    got = as_rows(image_load(tmp_path / "r.exr", flip_y=False, cast_to_linear=False))
    np.testing.assert_array_equal(got, fx.exr_expected(ch))


def test_exr_flip_channels_and_linear(tmp_path):
    ch = fx.sample_channels(6, 4, 3, fx.HALF, "RGB")
    fx.write_exr(str(tmp_path / "a.exr"), ch, fx.ZIP)
    want = fx.exr_expected(ch)                       # rows top first
    flipped = image_load(tmp_path / "a.exr")         # the reference's defaults: flip, no sRGB for .exr
    assert flipped.shape == (4, 6, 4)                # four channels (Magick's RGBA interface)
    np.testing.assert_array_equal(flipped[:, ::-1].transpose(1, 0, 2), want)   # y = 0 is the bottom row
    three = image_load(tmp_path / "a.exr", need_channel=3)
    np.testing.assert_array_equal(three, flipped[..., :3])
    five = image_load(tmp_path / "a.exr", need_channel=5)
    assert np.all(five[..., 4] == 0) and np.array_equal(five[..., :4], flipped)
    lin = image_load(tmp_path / "a.exr", cast_to_linear=True)
    v = flipped.astype(np.float32)
    ref = np.where(v <= np.float32(0.04045), v * np.float32(1.0 / 12.92),
                   np.power((v + np.float32(0.055)) * np.float32(1.0 / 1.055), np.float32(2.4)))
    np.testing.assert_allclose(lin, ref, rtol=1e-6, atol=0)   # powf and constant rounding: an ulp or two


def test_half_conversions_pinned():
    # float -> half rounding (nearest even) and half -> float against numpy
    from radiantium_amd.native import lib  # noqa: F401  (library loads)
    vals = np.array([0.0, -0.0, 1.0, 65504.0, 65519.99, 65520.0, 1e-8, 2.0 ** -25, 2.0 ** -24 * 1.5,
                     6.1e-5, 3.14159265, -2.5e-6, np.inf, -np.inf], np.float32)
    ch = {"R": (fx.FLOAT, vals.reshape(1, -1)), "G": (fx.FLOAT, -vals.reshape(1, -1)),
          "B": (fx.HALF, vals.reshape(1, -1))}
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "h.exr")
        fx.write_exr(p, ch, fx.NONE)
        got = as_rows(image_load(p, flip_y=False, cast_to_linear=False))
    want = fx.exr_expected(ch)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("w,h,top", [(17, 9, True), (5, 3, True), (300, 4, False)])
def test_hdr_decode(tmp_path, w, h, top):
    rng = np.random.default_rng(w * 31 + h)
    rgbe = rng.integers(0, 256, (h, w, 4)).astype(np.uint8)
    rgbe[:, : w // 2] = rgbe[:, :1]        # runs for the run-length coder
    rgbe[0, 0, 3] = 0                      # exponent 0: black
    p = tmp_path / "a.hdr"
    fx.write_hdr(str(p), rgbe, top_first=top)
    got = as_rows(image_load(p, flip_y=False))
    assert got.shape == (h, w, 3)
    np.testing.assert_array_equal(got.view(np.uint32), fx.hdr_expected(rgbe).view(np.uint32))
    assert np.all(got[0, 0] == 0)


@pytest.mark.parametrize("gray,little", [(False, True), (False, False), (True, True)])
def test_pfm_decode(tmp_path, gray, little):
    rng = np.random.default_rng(7)
    img = (rng.random((6, 9)) if gray else rng.random((6, 9, 3))).astype(np.float32) * 100
    p = tmp_path / "a.pfm"
    fx.write_pfm(str(p), img, little=little)
    got = as_rows(image_load(p, flip_y=False, cast_to_linear=False))
    np.testing.assert_array_equal(got, fx.pfm_expected(img))


def test_exr_save_roundtrip(tmp_path):
    rng = np.random.default_rng(11)
    cb = (rng.lognormal(0, 2, (23, 17, 3))).astype(np.float32)   # ColorBuffer [x, y, c]
    cb[0, 0] = [70000.0, -1.0, 0.0]
    for comp in (0, 3):
        p = tmp_path / f"o{comp}.exr"
        image_save_exr(p, cb, compression=comp)
        raw = open(p, "rb").read()
        assert struct.unpack("<I", raw[:4])[0] == 20000630 and raw[4] == 2
        assert b"compression\0compression\0" in raw and b"channels\0chlist\0" in raw
        back = image_load(p, flip_y=False, cast_to_linear=False)   # file rows = ColorBuffer rows (SaveOpenExr's k = i)
        assert back.shape == (23, 17, 4)
        # Magick: half(float(QuantumScale * (c * 65535f)))
        q = (cb * np.float32(65535.0)).astype(np.float32)
        h = np.asarray(np.float32((1.0 / 65535.0) * q.astype(np.float64)), np.float16).astype(np.float32)
        np.testing.assert_array_equal(back[..., :3], fx.via_quantum(h))
        assert np.all(back[..., 3] == fx.via_quantum(1.0))


def test_png_save_readback_with_pillow(tmp_path):
    from PIL import Image
    rng = np.random.default_rng(13)
    cb = rng.random((12, 7, 3)).astype(np.float32) * 1.2 - 0.1   # [x, y, c], some out of [0, 1]
    p = tmp_path / "o.png"
    image_save_png(p, cb)
    im = np.asarray(Image.open(p))
    assert im.shape == (7, 12, 4) and np.all(im[..., 3] == 255)
    q = np.clip(cb * np.float32(65535.0), 0, 65535).astype(np.float32)
    want = np.where(q <= 0, 0, np.where(q / np.float32(257.0) >= 255, 255,
                                        np.floor(q / np.float32(257.0) + np.float32(0.5)))).astype(np.uint8)
    # PNG row i is ColorBuffer row height - 1 - i
    np.testing.assert_array_equal(im[..., :3], want[:, ::-1].transpose(1, 0, 2))
    image_save_png(tmp_path / "s.png", cb, to_srgb=True)
    s = np.asarray(Image.open(tmp_path / "s.png"))
    v = cb
    srgb = np.where(v <= np.float32(0.0031308), np.float32(12.92) * v,
                    np.float32(1.055) * np.power(v, np.float32(1.0 / 2.4)) - np.float32(0.055)).astype(np.float32)
    qs = np.clip(np.nan_to_num(srgb, nan=0.0) * np.float32(65535.0), 0, 65535).astype(np.float32)
    ws = np.where(qs <= 0, 0, np.where(qs / np.float32(257.0) >= 255, 255,
                                       np.floor(qs / np.float32(257.0) + np.float32(0.5)))).astype(np.int32)
    assert np.abs(s[..., :3].astype(np.int32) - ws[:, ::-1].transpose(1, 0, 2)).max() <= 1   # pow rounding


def test_pfm_save_roundtrip(tmp_path):
    cb = np.random.default_rng(17).random((5, 8, 3)).astype(np.float32)
    p = tmp_path / "o.pfm"
    image_save_pfm(p, cb)
    raw = open(p, "rb").read()
    head = raw.split(b"\n", 3)
    assert head[0] == b"PF" and head[1] == b"5 8" and float(head[2]) < 0
    data = np.frombuffer(head[3], "<f4").reshape(8, 5, 3)    # first row = bottom = ColorBuffer row 0
    np.testing.assert_array_equal(data, cb.transpose(1, 0, 2))
    back = image_load(p, cast_to_linear=False)              # flip: y = 0 is the bottom row again
    np.testing.assert_array_equal(back, fx.via_quantum(cb))


def test_errors(tmp_path):
    with pytest.raises(RadiantiumError) as e:
        image_load(tmp_path / "missing.exr")
    assert e.value.code == abi.RD_ERR_IO
    (tmp_path / "junk.exr").write_bytes(b"not an image at all")
    with pytest.raises(RadiantiumError) as e:
        image_load(tmp_path / "junk.exr")
    assert e.value.code == abi.RD_ERR_UNSUPPORTED
    # a tiled OpenEXR file (version flag 0x200)
    (tmp_path / "t.exr").write_bytes(struct.pack("<II", 20000630, 2 | 0x200) + b"\0")
    with pytest.raises(RadiantiumError) as e:
        image_load(tmp_path / "t.exr")
    assert e.value.code == abi.RD_ERR_UNSUPPORTED
    # truncated PIZ data
    ch = fx.sample_channels(8, 8, 1, fx.HALF, "RGB")
    fx.write_exr(str(tmp_path / "p.exr"), ch, fx.PIZ)
    raw = open(tmp_path / "p.exr", "rb").read()
    (tmp_path / "p2.exr").write_bytes(raw[:-40])
    with pytest.raises(RadiantiumError) as e:
        image_load(tmp_path / "p2.exr")
    assert e.value.code == abi.RD_ERR_INVALID
