"""Display output of a rendered frame on the device, against the oracle.

* rh_read_display for every Quantize setting's type, and every other value
  type, equals the oracle's FormatBucketForDisplay of the oracle's channels
  and dither draws, byte for byte; uncropped and cropped frames.
* rh_display / rh_display_as with a crop window: the rectangles are the
  buckets' parts inside the crop window, in raster coordinates, in bucket
  order; rows span the crop window in scanline order; the pieces rebuild
  rh_read_display's image exactly once each.
* The file display driver's 16-bit and floating point files: the frame
  written through Hider.write_tiff holds rh_read_display's values of the
  unsigned type the driver asks for, with the crop window's origin and the
  full image size in its tags.
"""
import numpy as np
import pytest

from aqsis_amd import scenes
from aqsis_amd.hider import DSPY, DSPY_DTYPE, FILE_TYPE

pytestmark = pytest.mark.gpu

QUANTIZE = {
    "u8": (255.0, 0.0, 255.0, 0.5),
    "u16": (65535.0, 0.0, 65535.0, 0.5),
    "float": (0.0, 0.0, 0.0, 0.0),
    "s8": (127.0, -128.0, 127.0, 0.5),
    "s16": (32767.0, -32768.0, 32767.0, 1.0),
    "u32": (4294967295.0, 0.0, 4294967295.0, 0.0),
    "s32": (2147483647.0, -2147483648.0, 2147483647.0, 0.5),
}
CROPS = {"full": (0.0, 1.0, 0.0, 1.0), "crop": (0.13, 0.71, 0.2, 0.83)}


def _scene(q, crop, seed=61):
    opts, prims = scenes.patches_scene(seed=seed, n=60, res=(88, 56), samples=(2, 2))
    opts.quantize_one, opts.quantize_min, opts.quantize_max, opts.quantize_dither = q
    opts.crop[:] = list(crop)
    opts.exposure_gain = 1.3                  # values above 1: the clamps matter
    return opts, prims


@pytest.mark.parametrize("crop", list(CROPS))
@pytest.mark.parametrize("name", list(QUANTIZE))
def test_display_values_match_the_oracle(oracle, name, crop):
    from aqsis_amd.hider import Hider
    opts, prims = _scene(QUANTIZE[name], CROPS[crop])
    print(f"display {name} {crop}", flush=True)
    ref = oracle.render(opts, prims, occlusion_cull=True, samples=False)
    _, dither = oracle.sample_positions(opts)
    with Hider(opts) as h:
        h.post_surfaces(prims)
        h.render()
        assert h.display_type == oracle.display_type(opts)
        assert np.array_equal(h.channels().view(np.uint32), ref["channels"].view(np.uint32))
        for t in DSPY.values():
            want = oracle.display_values(opts, ref["channels"], dither, t)
            got = h.display_values(t)
            assert got.dtype == want.dtype
            assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), t
        if h.display_type == DSPY["uint8"]:
            assert np.array_equal(h.display_values(), h.rgba8())


def _crop_px(opts):
    f = np.float32
    return [min(max(int(np.ceil(f(r) * f(c))), 0), r)
            for r, c in ((opts.xres, opts.crop[0]), (opts.xres, opts.crop[1]),
                         (opts.yres, opts.crop[2]), (opts.yres, opts.crop[3]))]


@pytest.mark.parametrize("scanline_order", [False, True])
@pytest.mark.parametrize("name", ["u16", "float", "s8"])
def test_cropped_delivery(name, scanline_order):
    from aqsis_amd.hider import Hider
    opts, prims = _scene(QUANTIZE[name], CROPS["crop"], seed=62)
    cx0, cx1, cy0, cy1 = _crop_px(opts)
    calls = []
    with Hider(opts) as h:
        h.post_surfaces(prims)
        h.render()
        assert h.origin == (cx0, cy0) and (h.width, h.height) == (cx1 - cx0, cy1 - cy0)
        img = h.display_values()
        h.display(lambda *a: calls.append(a), scanline_order=scanline_order)
    dt = img.dtype
    seen = np.zeros(img.shape[:2], np.int32)
    rebuilt = np.zeros_like(img)
    for x0, x1, y0, y1, es, data in calls:
        assert es == 4 * dt.itemsize
        assert cx0 <= x0 < x1 <= cx1 and cy0 <= y0 < y1 <= cy1
        rebuilt[y0 - cy0:y1 - cy0, x0 - cx0:x1 - cx0] = np.frombuffer(data, dt).reshape(y1 - y0, x1 - x0, 4)
        seen[y0 - cy0:y1 - cy0, x0 - cx0:x1 - cx0] += 1
    assert np.all(seen == 1)
    assert np.array_equal(rebuilt.view(np.uint8), img.view(np.uint8))
    got = [c[:4] for c in calls]
    if scanline_order:
        assert got == [(cx0, cx1, y, y + 1) for y in range(cy0, cy1)]
    else:
        want = []
        for by in range(0, opts.yres, 16):
            for bx in range(0, opts.xres, 16):
                r = (max(bx, cx0), min(bx + 16, cx1), max(by, cy0), min(by + 16, cy1))
                if r[0] < r[1] and r[2] < r[3]:
                    want.append(r)
        assert got == want


@pytest.mark.parametrize("name,crop", [("u16", "crop"), ("float", "full"), ("s16", "crop"), ("u32", "full")])
def test_typed_file_output(tmp_path, name, crop):
    from aqsis_amd.hider import Hider
    from test_display import _tiff
    opts, prims = _scene(QUANTIZE[name], CROPS[crop], seed=63)
    path = tmp_path / "frame.tif"
    with Hider(opts) as h:
        h.post_surfaces(prims)
        h.render()
        t = FILE_TYPE[h.display_type]
        img = h.display_values(t)
        h.write_tiff(path, "none")
        origin, full = h.origin, h.full_size
        cropped = (h.width, h.height) != full
    tags, strips = _tiff(path)
    assert strips == img.tobytes()
    assert tags[258] == (8 * np.dtype(DSPY_DTYPE[t]).itemsize,) * 4
    if cropped:
        assert tags[286] == (origin[0], 1) and tags[287] == (origin[1], 1)
        assert tags[33300] == (full[0],) and tags[33301] == (full[1],)
    else:
        assert tags[33300] == (img.shape[1],) and tags[33301] == (img.shape[0],)


def test_rib_file_display_of_a_16_bit_frame(tmp_path):
    """A RIB with Quantize to 16 bits and a crop window writes the 16-bit
    cropped frame, the same values the ABI gives."""
    from aqsis_amd import rib
    from aqsis_amd.hider import Hider
    from test_display import _tiff
    opts, prims = _scene(QUANTIZE["u16"], CROPS["crop"], seed=64)
    path = tmp_path / "scene.rib"
    path.write_text('Display "+out.tif" "file" "rgba" "compression" "none"\n' + scenes.to_rib(opts, prims))
    written = rib.render_file(str(path))
    with Hider(opts) as h:
        h.post_surfaces(prims)
        h.render()
        img = h.display_values()
    tags, strips = _tiff(written[0])
    assert img.dtype == np.uint16 and strips == img.tobytes()


REQUESTS = {                       # mode, Quantize settings of the request
    "z_float": ("z", (0.0, 0.0, 0.0, 0.0)),
    "z_u16_dither": ("z", (1000.0, 0.0, 65535.0, 1.0)),
    "rgbaz_float": ("rgbaz", (0.0, 0.0, 0.0, 0.0)),
    "rgba_s16": ("rgba", (32767.0, -32768.0, 32767.0, 1.0)),
    "a_u8": ("a", (255.0, 0.0, 255.0, 0.5)),
    "rgb_u8": ("rgb", (255.0, 0.0, 255.0, 0.5)),
}


@pytest.mark.parametrize("index", [0, 1])
@pytest.mark.parametrize("req", list(REQUESTS))
def test_display_request_values_match_the_oracle(oracle, req, index):
    """rh_read_display_request: the frame's filtered channels as a display
    request of any mode receives them, with the request's own Quantize
    settings and the dither draws of its place among the requests (two
    requests here), against the oracle's restatement on the oracle's
    channels; a cropped frame."""
    from aqsis_amd.hider import Hider, mode_bits
    mode, q = REQUESTS[req]
    opts, prims = _scene(QUANTIZE["u8"], CROPS["crop"], seed=65)
    opts.num_displays = 2
    print(f"display request {req} index {index}", flush=True)
    ref = oracle.render(opts, prims, occlusion_cull=True, samples=False)
    _, dither = oracle.sample_positions(opts, display=index)
    with Hider(opts) as h:
        h.post_surfaces(prims)
        h.render()
        got = h.display_request_values(mode, q, index=index)
        assert np.array_equal(h.channels().view(np.uint32), ref["channels"].view(np.uint32))
    from aqsis_amd.hider import lib
    import ctypes as C
    t = C.c_int32()
    lib().rh_select_format(q[0], q[1], q[2], C.byref(t))
    want = oracle.display_request_values(q, mode_bits(mode), ref["channels"], dither, t.value)
    assert got.dtype == want.dtype and got.shape == want.shape
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
