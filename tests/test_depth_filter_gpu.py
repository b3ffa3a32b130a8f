"""Depth filters, the z display and matte surfaces on the device, against
the oracle, on the quad stacks whose oracle results the CPU tests pin by
closed forms: every filter with and without a z display, opaque, transparent
and matte alpha mattes, and the average filter's 0 / 0 samples (NaN: by
position, since the default NaN of an invalid operation is the hardware's).
The random scenes of every mode run in the parity tests."""
import numpy as np
import pytest

from aqsis_amd.abi import DEPTH_FILTERS
from test_depth_filter import QUADS, _opts, _quad

pytestmark = pytest.mark.gpu

STACKS = {
    "stack": [("T", {}), ("Q3", {}), ("Q1", {}), ("Q2", {})],
    "transparent_only": [("T", {})],
    "opaque_matte": [("Q1", {"matte": 1}), ("Q3", {})],
    "transparent_matte": [("T", {"matte": 1}), ("Q2", {}), ("Q3", {})],
    "matte_alpha": [("T", {"matte": 2}), ("Q1", {}), ("Q3", {})],
    "constant_matte": [("T", {"matte": 1, "shading_interpolation": 1}), ("Q1", {"matte": 1, "shading_interpolation": 1}),
                       ("Q3", {"shading_interpolation": 1})],
}


def _same(a, b):
    """Bit-equal, NaNs by position."""
    na, nb = np.isnan(a), np.isnan(b)
    assert np.array_equal(na, nb)
    assert np.array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32))


def _close(got, ref, tol=1e-5):
    """Filtered channels: finite values within tol, infinities bit-equal,
    NaNs by position."""
    fin, nan = np.isfinite(ref), np.isnan(ref)
    assert np.array_equal(fin, np.isfinite(got)) and np.array_equal(nan, np.isnan(got))
    inf = ~fin & ~nan
    assert np.array_equal(got[inf].view(np.uint32), ref[inf].view(np.uint32))
    d = np.abs(got[fin].astype(np.float64) - ref[fin].astype(np.float64))
    assert d.size == 0 or d.max() <= tol


@pytest.mark.parametrize("z", [0, 1])
@pytest.mark.parametrize("filt", list(DEPTH_FILTERS))
@pytest.mark.parametrize("stack", list(STACKS))
def test_stack_matches_the_oracle(oracle, stack, filt, z):
    from aqsis_amd.hider import Hider
    o = _opts(depth_filter=DEPTH_FILTERS[filt], display_z=z)
    prims = [_quad(o, n, **kw) for n, kw in STACKS[stack]]
    print(f"stack {stack} filter {filt} z {z}", flush=True)
    ref = oracle.render(o, prims, occlusion_cull=True)
    with Hider(o) as h:
        h.post_surfaces(prims)
        h.render()
        ids, depth, rgbaz = h.samples()
        ch = h.channels()
        st = h.stats()
    assert np.array_equal(ids, ref["mp_id"])
    _same(depth, ref["depth"])
    _same(rgbaz, ref["rgbaz"])
    _close(ch, ref["channels"])
    for k in ("spl_count", "spl_bound_hits", "spl_hits", "occlusion_culled"):
        assert st[k] == ref["stats"][k], k
