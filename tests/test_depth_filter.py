"""Depth filters, the z display mode and matte surfaces in the oracle (CPU),
pinned by closed forms over stacks of camera-facing quads.

The depth of a quad at a sample does not depend on the other surfaces, so
each stack is checked against renders of its quads alone (`_solo`):

* min (the default): the nearest opaque depth;
* midpoint with a z display: StoreSample keeps the nearest opaque hit and
  occludes with the second nearest (bucketprocessor.cpp:1499-1525); the
  combine gives (d0 + d1) * 0.5, FLT_MAX standing in for a missing second
  depth, and a list of one transparent hit FLT_MAX (imagepixel.cpp:265-275,
  319-326);
* max with a z display: no hit is cullable, every hit is listed, the depth
  is the farthest opaque one (imagepixel.cpp:275-278);
* average with a z display: the float mean, front to back, of the depths of
  the hits with any opacity channel at zthreshold, 0 / 0 without one
  (imagepixel.cpp:279-295);
* matte (RiMatte 1): an opaque matte sample without a list is transparent
  black; a listed matte hit lerps towards zero, the colour by the hit's
  opacity and the opacity by its colour (imagepixel.cpp:202-213, 308-318);
  matte alpha (RiMatte 2) is stored as an occluding hit whatever its
  opacity (bucketprocessor.cpp:1490-1491).
"""
import numpy as np
import pytest

from aqsis_amd import rib, scenes
from aqsis_amd.abi import DEPTH_FILTERS, FLT_MAX, default_options

F = np.float32
NONE = 0xFFFFFFFF


def _opts(**kw):
    return default_options(xres=48, yres=40, fov=45.0, xsamples=2, ysamples=2, filter="box", filter_xwidth=1.0,
                           filter_ywidth=1.0, **kw)


QUADS = {  # rect, depth, colour, opacity
    "T": ((10.0, 30.0, 10.0, 30.0), 3.0, (0.9, 0.2, 0.4), (0.5, 0.5, 0.5)),
    "Q1": ((4.0, 44.0, 4.0, 36.0), 5.0, (0.2, 0.7, 0.3), (1.0, 1.0, 1.0)),
    "Q2": ((8.0, 40.0, 8.0, 32.0), 7.0, (0.6, 0.1, 0.8), (1.0, 1.0, 1.0)),
    "Q3": ((2.0, 46.0, 2.0, 38.0), 9.0, (0.3, 0.3, 0.9), (1.0, 1.0, 1.0)),
}


def _quad(o, name, **kw):
    rect, z, cs, os_ = QUADS[name]
    p = scenes.raster_quad(o, rect, z, cs, os_)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _render(oracle, o, prims):
    r = oracle.render(o, prims, occlusion_cull=True)
    return r


def _solo(oracle, name, **kw):
    """Per-sample (covered, rgbaz, hit depth) of one quad rendered alone
    (the hit depth of the front-most hit: for the transparent quad, whose
    depth channel holds no opaque depth, its own)."""
    o = _opts()
    r = _render(oracle, o, [_quad(o, name, **kw)])
    return r["mp_id"] != NONE, r["rgbaz"], r["depth"]


@pytest.fixture(scope="module")
def solo(oracle):
    return {n: _solo(oracle, n) for n in QUADS}


def _stack(oracle, filt, z, names=("T", "Q3", "Q1", "Q2"), **kw):
    o = _opts(depth_filter=DEPTH_FILTERS[filt], display_z=z)
    return _render(oracle, o, [_quad(o, n, **kw) for n in names])


def _expected_depth(solo, filt, names=("T", "Q3", "Q1", "Q2")):
    """The depth per sample from the quads' solo depths and opacities.  The
    store: the opaque quads' nearest hit occludes (with the midpoint store
    the second nearest opaque depth is the occluding depth); the
    transparent quad's hit is listed (with max and average every hit is).
    The combine: back to front, a hit whose interpolated opacity reaches
    zthreshold (1: smooth interpolation of an opacity of 1 can give
    0.99999994, which does not) is an opaque depth."""
    n = solo["Q1"][0].size
    out = np.full(n, F(FLT_MAX), F)
    every = filt in ("max", "average")
    for i in range(n):
        hits = sorted((F(solo[m][2][i]), m) for m in names if solo[m][0][i])
        opaque = [h for h in hits if h[1] != "T"]
        if not hits:
            continue
        # hits the combine sees; the occluding depth it starts from
        if every:
            listed, occl_z = hits, F(FLT_MAX)
        else:
            listed = [h for h in hits if h[1] == "T"]
            if opaque:
                listed = sorted(listed + [opaque[0]])
            if filt == "midpoint":
                occl_z = opaque[1][0] if len(opaque) > 1 else F(FLT_MAX)
            else:
                occl_z = opaque[0][0] if opaque else F(FLT_MAX)
        if not every and not any(h[1] == "T" for h in hits):
            # no list: the occluding hit alone
            d = opaque[0][0]
            out[i] = F(0.5 * float(F(d + occl_z))) if filt == "midpoint" else d
            continue
        depth_of = {m: F(solo[m][2][i]) for _, m in listed}
        op0, op1, mx = occl_z, F(FLT_MAX), F(FLT_MAX)
        for _, m in reversed(listed):
            if np.all(solo[m][1][i, 3:6] >= F(1.0)):
                op0, op1 = depth_of[m], op0
                if not mx < F(FLT_MAX):
                    mx = depth_of[m]
        if filt == "midpoint":
            with np.errstate(over="ignore"):
                out[i] = F(F(op0 + op1) * F(0.5)) if len(listed) > 1 else F(FLT_MAX)
        elif filt == "max":
            out[i] = mx
        elif filt == "average":
            tot, cnt = F(0.0), 0
            for _, m in listed:
                if np.any(solo[m][1][i, 3:6] >= F(1.0)):
                    tot = F(tot + depth_of[m])
                    cnt += 1
            with np.errstate(invalid="ignore", divide="ignore"):
                out[i] = F(tot / F(cnt))
        else:
            out[i] = op0
    return out


@pytest.mark.parametrize("filt", ["min", "midpoint", "max", "average"])
def test_depth_filter_closed_forms(oracle, solo, filt):
    r = _stack(oracle, filt, 1)
    got = r["rgbaz"][:, 6]
    want = _expected_depth(solo, filt)
    covered = np.zeros_like(got, bool)
    for n in QUADS:
        covered |= solo[n][0]
    assert np.array_equal(r["mp_id"] != NONE, covered)
    # FLT_MAX plus a depth rounds to FLT_MAX: one opaque depth alone
    # midpoints to FLT_MAX / 2
    assert np.array_equal(got[covered].view(np.uint32), want[covered].view(np.uint32))
    if filt == "midpoint":
        only_q3 = solo["Q3"][0] & ~solo["Q1"][0] & ~solo["T"][0]
        assert only_q3.any() and np.all(got[only_q3] == F(FLT_MAX) * F(0.5))


def test_midpoint_single_opaque_without_list_is_half_flt_max(oracle, solo):
    # no transparent surface: the combine's no-list branch,
    # 0.5 * (double)(d + occlZ) with occlZ = FLT_MAX
    r = _stack(oracle, "midpoint", 1, names=("Q3",))
    got = r["rgbaz"][solo["Q3"][0], 6]
    assert np.all(got == F(0.5 * float(F(FLT_MAX))))


def test_midpoint_without_z_display_keeps_the_nearest_depth(oracle, solo):
    # without a z display the store is the usual one (occlZ = the occluding
    # depth): no-list samples keep their depth, listed ones average the
    # nearest opaque depth with FLT_MAX
    r = _stack(oracle, "midpoint", 0, names=("Q3", "Q1", "Q2"))
    got = r["rgbaz"][:, 6]
    want = _expected_depth(solo, "min", names=("Q3", "Q1", "Q2"))
    cov = solo["Q3"][0]
    assert np.array_equal(got[cov].view(np.uint32), want[cov].view(np.uint32))


def test_max_and_average_composite_every_hit_like_min(oracle):
    """With every hit listed, opaque layers behind reset the composite: the
    colour and opacity equal the default filter's (constant interpolation,
    so that opaque hits have opacity exactly 1); nothing is culled."""
    base = _stack(oracle, "min", 1, shading_interpolation=1)
    for filt in ("max", "average"):
        r = _stack(oracle, filt, 1, shading_interpolation=1)
        assert np.array_equal(r["rgbaz"][:, :6].view(np.uint32), base["rgbaz"][:, :6].view(np.uint32))
        assert r["stats"]["occlusion_culled"] == 0
        # the hidden layers are hit and stored: more stored hits than min
        assert r["stats"]["spl_hits"] > base["stats"]["spl_hits"]


def test_average_zero_terms_is_nan(oracle, solo):
    # only the transparent quad (opacity 0.5 < zthreshold 1): 0 / 0
    r = _stack(oracle, "average", 1, names=("T",))
    got = r["rgbaz"][solo["T"][0], 6]
    assert np.all(np.isnan(got))


def test_average_counts_any_channel_at_the_threshold(oracle):
    # opacity (0.7, 0.2, 0.2) against zthreshold 0.6: counted by the mean
    # (any channel) but not an opaque depth for the composite (all channels)
    o = _opts(depth_filter=DEPTH_FILTERS["average"], display_z=1)
    o.zthreshold[:] = [0.6, 0.6, 0.6]
    a = scenes.raster_quad(o, (4.0, 44.0, 4.0, 36.0), 5.0, (0.2, 0.7, 0.3), (0.7, 0.2, 0.2))
    b = scenes.raster_quad(o, (4.0, 44.0, 4.0, 36.0), 8.0, (0.6, 0.1, 0.8), (0.1, 0.1, 0.1))
    r = _render(oracle, o, [a, b])
    sa = _render(oracle, o, [a])["rgbaz"][:, 6]
    cov = r["mp_id"] != NONE
    assert np.array_equal(r["rgbaz"][cov, 6].view(np.uint32), (F(0) + sa[cov]).view(np.uint32))
    o.depth_filter = DEPTH_FILTERS["max"]
    r = _render(oracle, o, [a, b])
    assert np.all(r["rgbaz"][cov, 6] == F(FLT_MAX))     # no opaque depth


def test_opaque_matte_is_transparent_black(oracle, solo):
    o = _opts()
    r = _render(oracle, o, [_quad(o, "Q1", matte=1), _quad(o, "Q3")])
    m = solo["Q1"][0]
    assert np.all(r["rgbaz"][m, :6] == 0.0)
    assert np.array_equal(r["rgbaz"][m, 6].view(np.uint32), solo["Q1"][1][m, 6].view(np.uint32))
    rest = solo["Q3"][0] & ~m
    assert np.array_equal(r["rgbaz"][rest].view(np.uint32), solo["Q3"][1][rest].view(np.uint32))
    # the filtered alpha is zero inside the matte
    ch = r["channels"]
    assert ch[10:30, 10:30, 6].max() == 0.0


def test_transparent_matte_lerps(oracle, solo):
    o = _opts()
    r = _render(oracle, o, [_quad(o, "T", matte=1), _quad(o, "Q3")])
    t = solo["T"][0]
    Ci, Oi = solo["T"][1][t, 0:3], solo["T"][1][t, 3:6]
    Cb, Ob = solo["Q3"][1][t, 0:3], solo["Q3"][1][t, 3:6]
    one, zero = F(1), F(0)
    col = (one - Oi) * Cb + Oi * zero
    opa = (one - Ci) * Ob + Ci * zero
    assert np.array_equal(r["rgbaz"][t, 0:3].view(np.uint32), col.astype(F).view(np.uint32))
    assert np.array_equal(r["rgbaz"][t, 3:6].view(np.uint32), opa.astype(F).view(np.uint32))


def test_listed_opaque_matte_keeps_one_minus_colour_opacity(oracle, solo):
    # every hit listed (max with a z display): the opaque matte goes through
    # the lerps, colour (1 - 1) c + 1 * 0 = 0, opacity (1 - Ci) * 1 + Ci * 0
    # (constant interpolation: opacities exactly 1, colours exactly Cs)
    o = _opts(depth_filter=DEPTH_FILTERS["max"], display_z=1)
    r = _render(oracle, o, [_quad(o, "Q1", matte=1, shading_interpolation=1),
                            _quad(o, "Q3", shading_interpolation=1)])
    m = solo["Q1"][0]
    Ci = np.broadcast_to(np.array(QUADS["Q1"][2], F), (int(m.sum()), 3))
    assert np.all(r["rgbaz"][m, 0:3] == 0.0)
    want = (F(1) - Ci) * F(1) + Ci * F(0)
    assert np.array_equal(r["rgbaz"][m, 3:6].view(np.uint32), want.astype(F).view(np.uint32))


def test_matte_alpha_occludes_whatever_its_opacity(oracle, solo):
    o = _opts()
    r = _render(oracle, o, [_quad(o, "T", matte=2), _quad(o, "Q3")])
    t = solo["T"][0]
    # the matte alpha hit alone: its own colour and opacity and, as the
    # occluding hit, its depth; Q3 behind it rejected
    assert np.array_equal(r["rgbaz"][t, :6].view(np.uint32), solo["T"][1][t, :6].view(np.uint32))
    assert np.array_equal(r["rgbaz"][t, 6].view(np.uint32), solo["T"][2][t].view(np.uint32))
    r2 = _render(oracle, o, [_quad(o, "T"), _quad(o, "Q3")])
    assert not np.array_equal(r2["rgbaz"][t, :6].view(np.uint32), solo["T"][1][t, :6].view(np.uint32))
    assert r["stats"]["spl_hits"] < r2["stats"]["spl_hits"]


# --- the RIB front end -------------------------------------------------------

def _parse(body):
    return rib.parse(body + "\nWorldBegin\nSurface \"constant\"\nMatte 2\nPatch \"bilinear\" \"P\" "
                     "[-1 -1 5 1 -1 5 -1 1 5 1 1 5]\nAttributeBegin\nMatte 7\nSphere 1 -1 1 360\nAttributeEnd\n"
                     "Sphere 1 -1 1 360\nMatte 0\nSphere 1 -1 1 360\nWorldEnd\n")


def test_rib_depthfilter_display_z_and_matte():
    o, prims = _parse('Hider "hidden" "depthfilter" "max"\nDisplay "a.tif" "file" "rgbaz"')
    assert o.depth_filter == DEPTH_FILTERS["max"] and o.display_z == 1
    assert [p.matte for p in prims] == [2, 1, 2, 0]
    o, _ = _parse('Display "a.tif" "file" "z"\nDisplay "b.tif" "file" "rgba"')
    assert o.display_z == 0                   # replaced by the second request
    o, _ = _parse('Display "a.tif" "file" "rgba"\nDisplay "+b.tif" "file" "z"')
    assert o.display_z == 1 and o.num_displays == 2
    o, _ = _parse('Display "a.tif" "file" "rgbz"')
    assert o.display_z == 1
    o, _ = _parse('Display "a.tif" "file" "depth"')
    assert o.display_z == 0
    with pytest.warns(UserWarning, match="depthfilter"):
        o, _ = _parse('Hider "hidden" "depthfilter" "median"')
    assert o.depth_filter == DEPTH_FILTERS["min"]


def test_invalid_values_refused():
    from aqsis_amd.hider import HiderError, estimate
    o = _opts(depth_filter=4)
    with pytest.raises(HiderError, match="depth filter"):
        estimate(o, [_quad(o, "Q1")])
    o = _opts()
    with pytest.raises(HiderError, match="matte"):
        estimate(o, [_quad(o, "Q1", matte=3)])
