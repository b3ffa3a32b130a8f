"""HIP hider vs the oracle on the same scenes, through the C ABI.

The bar: per-sample winning micropolygon id, depth and combined colour
bit-exact; filtered channels within 1e-5 absolute (they come out
bit-identical); 8-bit output identical; the hider counters SPL_count,
SPL_bound_hits and SPL_hits and the number of occlusion-culled surfaces
equal to the oracle's.  Both run the reference's default, occlusion culling
on for every surface whose Attribute "cull" "hidden" is 1: culling defers
surfaces to later buckets, which decides the winners at equal depths
(coplanar and decal cases).
"""
import numpy as np
import pytest

from aqsis_amd import scenes
from parity_cases import CASES

pytestmark = pytest.mark.gpu

PIXEL_TOL = 1e-5


def _render(opts, prims):
    from aqsis_amd.hider import Hider
    with Hider(opts) as h:
        h.post_surfaces(prims)
        h.render()
        ids, depth, rgbaz = h.samples()
        return {"channels": h.channels(), "rgba8": h.rgba8(), "mp_id": ids, "depth": depth,
                "rgbaz": rgbaz, "stats": h.stats()}


@pytest.mark.parametrize("case", list(CASES))
def test_parity(case, oracle):
    opts, prims = CASES[case]()
    print(f"case {case}: {opts.xres}x{opts.yres} spp {opts.xsamples}x{opts.ysamples} prims {len(prims)}",
          flush=True)
    ref = oracle.render(opts, prims, occlusion_cull=True)
    got = _render(opts, prims)
    assert np.array_equal(got["mp_id"], ref["mp_id"]), \
        f"winner id mismatch at {np.count_nonzero(got['mp_id'] != ref['mp_id'])} samples"
    assert np.array_equal(got["depth"].view(np.uint32), ref["depth"].view(np.uint32))
    assert np.array_equal(got["rgbaz"].view(np.uint32), ref["rgbaz"].view(np.uint32))
    # Non-finite values (a depth sum of FLT_MAX samples overflows) must match:
    # infinities bit for bit, NaNs (inf - inf where a filter has negative
    # lobes) by position only, since the default NaN of an invalid operation
    # is the hardware's (x86: sign bit set; the GPU: clear).  Finite values
    # within the tolerance.
    fin = np.isfinite(ref["channels"])
    nan = np.isnan(ref["channels"])
    assert np.array_equal(fin, np.isfinite(got["channels"]))
    assert np.array_equal(nan, np.isnan(got["channels"]))
    inf = ~fin & ~nan
    assert np.array_equal(got["channels"][inf].view(np.uint32), ref["channels"][inf].view(np.uint32))
    diff = np.abs(got["channels"][fin].astype(np.float64) - ref["channels"][fin].astype(np.float64))
    assert diff.size == 0 or diff.max() <= PIXEL_TOL
    assert np.array_equal(got["rgba8"], ref["rgba8"])
    for k in ("spl_count", "spl_bound_hits", "spl_hits", "inside", "inside_nonlinear", "grids", "micropolygons",
              "occlusion_culled"):
        assert got["stats"][k] == ref["stats"][k], k
    hit = np.count_nonzero(ref["mp_id"] != 0xFFFFFFFF)
    assert hit > 0.05 * ref["mp_id"].size  # the scene covers the frame


def test_rerender_identical(oracle):
    """rh_render_device on resident inputs reproduces the first frame."""
    from aqsis_amd.hider import Hider
    opts, prims = CASES["patches_small"]()
    with Hider(opts) as h:
        h.post_surfaces(prims)
        h.render()
        a = (h.channels(), h.rgba8(), h.samples()[0])
        for _ in range(3):
            h.render_device()
        b = (h.channels(), h.rgba8(), h.samples()[0])
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.gpu
def test_contexts_of_one_shape_render_alike(oracle):
    """A second context of a frame shape takes its sample plan from the
    process's plan cache; it renders exactly what the first did and what the
    oracle draws for that shape."""
    from aqsis_amd.hider import Hider
    opts, prims = CASES["patches_transparent"]()
    outs = []
    for _ in range(2):
        with Hider(opts) as h:
            h.post_surfaces(prims)
            h.render()
            outs.append((h.channels(), h.rgba8(), h.samples()[0]))
    for x, y in zip(*outs):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    ref = oracle.render(opts, prims, occlusion_cull=True)
    assert np.array_equal(outs[1][2], ref["mp_id"])


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["no_surfaces", "all_outside"])
def test_frames_without_visible_surfaces(kind, oracle):
    """A frame with no surfaces, and one whose surfaces all lie outside the
    view, render the empty image: no hits, empty samples, channels and 8-bit
    output as the oracle's."""
    from aqsis_amd.hider import Hider
    opts, prims = scenes.patches_scene(seed=41, n=20, res=(64, 48), samples=(2, 2))
    if kind == "no_surfaces":
        prims = []
    else:
        for p in prims:
            for i in range(16):
                p.P[i][0] += 1000.0           # far to the side of the frustum
    ref = oracle.render(opts, prims, occlusion_cull=True)
    with Hider(opts) as h:
        h.post_surfaces(prims)
        h.render()
        ids, depth, _ = h.samples()
        st = h.stats()
        assert np.array_equal(h.channels().view(np.uint32), ref["channels"].view(np.uint32))
        assert np.array_equal(h.rgba8(), ref["rgba8"])
    assert np.array_equal(ids, ref["mp_id"]) and np.all(ids == 0xffffffff)
    assert st["spl_hits"] == ref["stats"]["spl_hits"] == 0
    assert st["spl_count"] == ref["stats"]["spl_count"]


@pytest.mark.gpu
@pytest.mark.parametrize("scanline_order", [False, True])
def test_display_delivery(scanline_order):
    """rh_display hands the 8-bit image to a display driver as
    CqDisplayRequest::DisplayBucket does: bucket rectangles in the render's
    bucket order (16x16, clipped at the image edge), or full-width rows in
    order once each row of buckets is complete; the pieces rebuild the
    image exactly."""
    from aqsis_amd.hider import Hider
    opts, prims = scenes.patches_scene(seed=43, n=40, res=(72, 40), samples=(2, 2))
    calls = []
    with Hider(opts) as h:
        h.post_surfaces(prims)
        h.render()
        img = h.rgba8()
        h.display(lambda *a: calls.append(a), scanline_order=scanline_order)
    rebuilt = np.zeros_like(img)
    for x0, x1, y0, y1, es, data in calls:
        assert es == 4
        rebuilt[y0:y1, x0:x1] = np.frombuffer(data, np.uint8).reshape(y1 - y0, x1 - x0, 4)
    assert np.array_equal(rebuilt, img)
    if scanline_order:
        assert [(c[0], c[1], c[2], c[3]) for c in calls] == [(0, 72, y, y + 1) for y in range(40)]
    else:
        want = [(bx, min(bx + 16, 72), by, min(by + 16, 40)) for by in range(0, 40, 16) for bx in range(0, 72, 16)]
        assert [(c[0], c[1], c[2], c[3]) for c in calls] == want


@pytest.mark.gpu
@pytest.mark.parametrize("compression", ["none", "lzw"])
def test_tiff_file_of_frame(tmp_path, compression):
    """The frame written as the file display driver writes it (the driver's
    buffer filled from rh_display, then rh_tiff_write) holds the frame's
    8-bit rgba."""
    from aqsis_amd.hider import Hider
    from test_tiff import _raw_pixels, _read, _same
    opts, prims = scenes.patches_scene(seed=44, n=60, res=(80, 48), samples=(2, 2))
    path = tmp_path / "frame.tif"
    with Hider(opts) as h:
        h.post_surfaces(prims)
        h.render()
        img = h.rgba8()
        h.write_tiff(path, compression)
    got, tags = _read(path)
    assert tags[256] == 80 and tags[257] == 48
    assert _same(got, img)
    if compression == "none":
        assert _raw_pixels(path) == img.tobytes()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["rgba", "rgb"])
def test_rib_file_display(tmp_path, mode):
    """A RIB with a "file" display renders to that TIFF file: the pixels of
    the frame the same scene gives through the ABI."""
    from aqsis_amd import rib
    from aqsis_amd.hider import Hider
    from test_tiff import _read, _same
    opts, prims = scenes.patches_scene(seed=45, n=50, res=(64, 40), samples=(2, 2))
    path = tmp_path / "scene.rib"
    path.write_text(f'Display "+out.tif" "file" "{mode}"\n' + scenes.to_rib(opts, prims))
    written = rib.render_file(str(path))
    assert written == [str(tmp_path / "out.tif")]
    with Hider(opts) as h:
        h.post_surfaces(prims)
        h.render()
        img = h.rgba8()[..., :len(mode)]
    got, tags = _read(written[0])
    assert tags[259] == 5 and tags[277] == len(mode)
    assert _same(got, img)
