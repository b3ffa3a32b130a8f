"""8-bit TIFF output (rh_tiff_write), the file display driver's integer
branch (tools/displays/file/display.cpp:227-350).  Files are read back with
Pillow's TIFF decoder, an implementation independent of the writer, so a
wrong LZW code width, table reset, strip layout or tag fails here.
Pillow un-premultiplies associated alpha when it decodes, so rgba is
compared channel for channel where alpha is 255 (there the conversion is
the identity) and alpha everywhere; uncompressed files are also parsed
byte for byte by the small reader below."""
import struct

import numpy as np
import pytest
from PIL import Image

from aqsis_amd.hider import HiderError, write_tiff


def _read(path):
    with Image.open(path) as im:
        tags = dict(im.tag_v2)  # before load(), which drops Orientation
        im.load()
        return np.asarray(im), tags


def _raw_pixels(path):
    """Strips of an uncompressed little-endian TIFF, concatenated."""
    b = open(path, "rb").read()
    assert b[:4] == b"II*\0"
    ifd = struct.unpack_from("<I", b, 4)[0]
    n = struct.unpack_from("<H", b, ifd)[0]
    tags = {}
    for k in range(n):
        tag, typ, cnt, val = struct.unpack_from("<HHII", b, ifd + 2 + 12 * k)
        if typ == 4 and cnt > 1:
            val = struct.unpack_from(f"<{cnt}I", b, val)
        elif typ == 4:
            val = (val,)
        tags[tag] = val
    return b"".join(b[o:o + c] for o, c in zip(tags[273], tags[279]))


def _same(got, img):
    if img.shape[2] == 3:
        return np.array_equal(got, img)
    opaque = img[..., 3] == 255
    return (np.array_equal(got[..., 3], img[..., 3]) and opaque.any()
            and np.array_equal(got[opaque], img[opaque]))


def _opaque_half(a, rng):
    a[..., 3] = np.where(rng.random(a.shape[:2]) < 0.5, 255, a[..., 3])
    return a


def _images():
    rng = np.random.default_rng(5)
    yield "noise_rgba", _opaque_half(rng.integers(0, 256, (61, 47, 4), dtype=np.uint8), rng)
    yield "noise_rgb", rng.integers(0, 256, (33, 129, 3), dtype=np.uint8)
    # Long runs: strings grow past 12-bit codes and the table resets.
    y, x = np.mgrid[0:300, 0:500]
    yield "gradient_rgba", np.stack([x % 256, y % 256, (x // 7) % 256, np.full_like(x, 255)], -1).astype(np.uint8)
    yield "flat_rgba", np.dstack([np.full((200, 300, 3), 17, np.uint8), np.full((200, 300), 255, np.uint8)])
    # A row wider than a default strip (8 KiB): one row per strip.
    yield "wide_rgba", _opaque_half(rng.integers(0, 4, (5, 2100, 4), dtype=np.uint8), rng)
    yield "single_pixel", np.array([[[1, 2, 3, 255]]], np.uint8)


@pytest.mark.parametrize("compression", ["none", "lzw"])
@pytest.mark.parametrize("name,img", list(_images()), ids=[n for n, _ in _images()])
def test_round_trip(tmp_path, compression, name, img):
    path = tmp_path / f"{name}.tif"
    write_tiff(path, img, compression)
    got, tags = _read(path)
    assert got.shape == img.shape
    assert _same(got, img)
    if compression == "none":
        assert _raw_pixels(path) == img.tobytes()
    h, w, c = img.shape
    assert tags[259] == {"none": 1, "lzw": 5}[compression]
    assert tags[262] == 2                      # RGB
    assert tags[274] == 1                      # top-left
    assert tags[277] == c
    assert tags[278] == max(1, 8192 // (w * c))  # libtiff's default strip height
    assert tags[296] == 1                      # resolution unit none
    assert float(tags[282]) == 1.0 and float(tags[283]) == 1.0
    assert (338 in tags) == (c == 4)
    if c == 4:
        assert tags[338] == (1,)               # associated alpha
    assert tags[33300] == w and tags[33301] == h


def test_lzw_compresses_runs(tmp_path):
    img = np.full((256, 256, 4), 200, np.uint8)
    write_tiff(tmp_path / "a.tif", img, "lzw")
    write_tiff(tmp_path / "b.tif", img, "none")
    assert (tmp_path / "a.tif").stat().st_size * 20 < (tmp_path / "b.tif").stat().st_size


def test_origin_and_full_size(tmp_path):
    img = np.zeros((10, 12, 4), np.uint8)
    write_tiff(tmp_path / "c.tif", img, "lzw", origin=(32, 16), full_size=(640, 480), description="crop")
    _, tags = _read(tmp_path / "c.tif")
    assert float(tags[286]) == 32.0 and float(tags[287]) == 16.0
    assert tags[33300] == 640 and tags[33301] == 480
    assert tags[270] == "crop"


@pytest.mark.parametrize("shape,compression", [((4, 4, 2), "lzw"), ((4, 4, 5), "none"), ((0, 4, 4), "lzw")])
def test_refused(tmp_path, shape, compression):
    with pytest.raises(HiderError):
        write_tiff(tmp_path / "x.tif", np.zeros(shape, np.uint8), compression)


def test_unwritable_path(tmp_path):
    with pytest.raises(HiderError, match="cannot open"):
        write_tiff(tmp_path / "no_such_dir" / "x.tif", np.zeros((2, 2, 4), np.uint8))
