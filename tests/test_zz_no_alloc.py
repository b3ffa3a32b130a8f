"""rh_render_device_noalloc (the no-allocation form of rh_render_device,
for callers that cannot allocate) against rh_render_device, through the C
ABI: per-sample ids, depths and combined colours, channels, 8-bit output and
the hider counters bit-identical, on three calls in a row with smaller
frames each time (transparent patches, depth of field, moving patches), the
later ones finding the shared scratch array holding the earlier frame."""
import numpy as np
import pytest

from parity_cases import CASES

pytestmark = pytest.mark.gpu

COUNTERS = ("spl_count", "spl_bound_hits", "spl_hits", "inside", "inside_nonlinear", "grids", "micropolygons",
            "mp_bucket_refs", "occlusion_culled")


def _frame(h):
    ids, depth, rgbaz = h.samples()
    st = h.stats()
    return {"ids": ids, "depth": depth, "rgbaz": rgbaz, "channels": h.channels(), "rgba8": h.rgba8(),
            "stats": {k: st[k] for k in COUNTERS}}


def test_noalloc_frames_equal_render_device():
    from aqsis_amd.hider import Hider
    last = None
    for case in ("patches_transparent", "patches_dof", "motion_transparent_8x8"):
        opts, prims = CASES[case]()
        size = opts.xres * opts.yres
        assert last is None or size < last
        last = size
        print(f"no-alloc case {case}: {opts.xres}x{opts.yres}", flush=True)
        with Hider(opts) as h:
            h.post_surfaces(prims)
            h.render()
            h.render_device()
            ref = _frame(h)
            h.render_device_noalloc()
            got = _frame(h)
        for k in ("ids", "rgba8"):
            assert np.array_equal(got[k], ref[k]), (case, k)
        for k in ("depth", "rgbaz", "channels"):
            assert np.array_equal(got[k].view(np.uint32), ref[k].view(np.uint32)), (case, k)
        assert got["stats"] == ref["stats"], case
        assert ref["stats"]["spl_hits"] > 0
