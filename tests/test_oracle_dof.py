"""Depth of field in the oracle, pinned by closed forms (CPU).

The reference's lens (RiDepthOfField fstop focallength focaldistance;
CqRenderer::SetDepthOfFieldData / GetCircleOfConfusion, renderer.h:368-406;
the CoC scale of options.cpp:160-171) is a thin lens: a point at camera
depth z blurs into a disk of raster radius
    0.5 * (L / N) * f / (f + L / N) * |1/z - 1/f| * (raster units per camera
    unit at depth 1),
and a sample with lens offset o in the unit disk sees the scene shifted by
CoC(z) * o (micropolygon.cpp:1561-1593).  These tests hold the oracle's
restatement to that, written here from the optics, not from the oracle:
    * the circle of confusion and MinCoCForBound against the formula;
    * lens offsets: projectToCircle of each pixel's pattern, inside the unit
      disk, each in the bound of its lens cell (CalculateDofBounds);
    * a scene on the focal plane renders exactly as the pinhole camera;
    * a defocused opaque quad parallel to the image plane is hit exactly at
      the samples s with s + CoC(z) * o inside its raster rectangle;
    * the dicing rate shrinks with the CoC area (AdjustedShadingRate).
"""
import math

import numpy as np
import pytest

from aqsis_amd import scenes
from aqsis_amd.abi import default_options


def _dof_opts(**kw):
    base = dict(xres=64, yres=48, fov=45.0, xsamples=3, ysamples=3, filter="box", filter_xwidth=1.0,
                filter_ywidth=1.0)
    base.update(kw)
    return default_options(**base)


def _raster_per_unit(o):
    """Raster pixels per camera unit at depth 1, along x and y (perspective,
    pixel aspect 1): the screen window spans 2 * max(aspect, 1) * tan(fov/2)
    at depth 1 across xres pixels."""
    aspect = o.xres / o.yres
    t = math.tan(math.radians(o.fov) / 2.0)
    return o.xres / (2.0 * t * max(aspect, 1.0)), o.yres / (2.0 * t * max(1.0 / aspect, 1.0))


def _coc(o, z):
    N, L, f = (float(v) for v in o.dof)
    D = L / N
    kx, ky = _raster_per_unit(o)
    c = 0.5 * D * f / (f + D) * abs(1.0 / z - 1.0 / f)
    return c * kx, c * ky


def test_circle_of_confusion_closed_form(oracle):
    o = _dof_opts()
    o.dof[:] = [2.8, 1.5, 10.0]
    z = np.array([2.0, 5.0, 9.0, 10.0, 11.0, 30.0, 200.0], np.float32)
    coc, _, mn = oracle.dof(o, z)
    for i, zi in enumerate(z):
        ex = _coc(o, float(zi))
        assert coc[i] == pytest.approx(ex, rel=2e-6, abs=1e-7)
    assert coc[3].tolist() == [0.0, 0.0]            # on the focal plane
    # MinCoCForBound: 0 when [z0, z1] spans the focal plane, else the nearer end
    for i in range(len(z) - 1):
        z0, z1 = float(z[i]), float(z[i + 1])
        if (z0 - 10.0) * (z1 - 10.0) < 0:
            assert mn[i] == 0.0
        else:
            ex = min(min(_coc(o, z0)), min(_coc(o, z1)))
            assert mn[i] == pytest.approx(ex, rel=2e-6, abs=1e-7)


def _project_to_circle(p):
    r = np.hypot(p[..., 0], p[..., 1])
    m = np.maximum(np.abs(p[..., 0]), np.abs(p[..., 1]))
    adj = np.where(r > 0, m / np.where(r > 0, r, 1.0), 0.0)
    return p * adj[..., None]


@pytest.mark.parametrize("nx,ny", [(3, 3), (4, 2), (1, 1)])
def test_lens_offsets_are_the_pixel_patterns_on_the_disk(oracle, nx, ny):
    """Each pixel's lens offsets are projectToCircle(2p - 1) of one of the
    250 multijitter patterns, offset k of the pattern going to the sample
    visited through lens cell k; all lie in the unit disk.  A lens cell's
    bound holds the projected corners of its square of [-1, 1]^2."""
    o = _dof_opts(xres=24, yres=16)
    o.xsamples, o.ysamples = nx, ny
    o.dof[:] = [4.0, 1.0, 10.0]
    ns = nx * ny
    lens, cell = oracle.sample_lens(o)
    lens = lens.reshape(-1, ns, 2).astype(np.float64)
    cell = cell.reshape(-1, ns)
    assert np.array_equal(np.sort(cell, axis=1), np.tile(np.arange(ns), (cell.shape[0], 1)))
    pos2d = np.zeros(250 * ns * 2, np.float32)
    pos1d = np.zeros(250 * ns, np.float32)
    shuf = np.zeros(250 * ns, np.int32)
    oracle.lib().oracle_jitter_tables(nx, ny, oracle._ptr(pos2d), oracle._ptr(pos1d), oracle._ptr(shuf))
    pats = _project_to_circle(2.0 * pos2d.reshape(250, ns, 2).astype(np.float64) - 1.0)
    assert np.all(np.hypot(lens[..., 0], lens[..., 1]) <= 1.0 + 1e-6)
    # the offsets in lens-cell order: pattern entry k is the sample of cell k
    by_cell = np.zeros_like(lens)
    np.put_along_axis(by_cell, cell[..., None], lens, axis=1)
    err = np.abs(by_cell[:, None] - pats[None]).max(axis=(2, 3))   # pixels x patterns
    assert np.all(err.min(axis=1) < 1e-6)
    # A cell clear of both axes is bounded by its projected corners; a cell
    # across an axis keeps its square's extent across that axis (projection
    # would pull the edge on the axis inwards).
    _, cells, _ = oracle.dof(o, np.array([5.0], np.float32))
    assert np.all(np.abs(cells) <= 1.0 + 1e-6)
    for k in range(ns):
        ix, iy = k % nx, k // nx
        x0, y0 = -1.0 + 2.0 * ix / nx, -1.0 + 2.0 * iy / ny
        x1, y1 = x0 + 2.0 / nx, y0 + 2.0 / ny
        corners = _project_to_circle(np.array([[x0, y0], [x1, y0], [x0, y1], [x1, y1]]))
        across_x, across_y = x0 < 0 < x1, y0 < 0 < y1
        if not across_x and not across_y:
            assert np.allclose(cells[k], [*corners.min(axis=0), *corners.max(axis=0)], atol=1e-6)
        if across_x:
            assert np.allclose(cells[k, [1, 3]], [y0, y1], atol=1e-6)
        if across_y:
            assert np.allclose(cells[k, [0, 2]], [x0, x1], atol=1e-6)


def _quads(o, rng, n, z):
    prims = []
    for _ in range(n):
        x0, y0 = rng.uniform(-4, o.xres - 4), rng.uniform(-4, o.yres - 4)
        w, h = rng.uniform(5, 20), rng.uniform(5, 20)
        prims.append(scenes.raster_quad(o, (x0, x0 + w, y0, y0 + h), z, tuple(rng.uniform(0, 1, 3))))
    return prims


def test_scene_on_the_focal_plane_renders_as_the_pinhole_camera(oracle):
    o = _dof_opts(filter="gaussian", filter_xwidth=2.0, filter_ywidth=2.0)
    prims = _quads(o, np.random.default_rng(3), 12, 10.0)
    pin = oracle.render(o, prims, occlusion_cull=True)
    o.dof[:] = [2.8, 1.0, 10.0]
    lens = oracle.render(o, prims, occlusion_cull=True)
    assert np.array_equal(pin["mp_id"], lens["mp_id"])
    assert np.array_equal(pin["depth"].view(np.uint32), lens["depth"].view(np.uint32))
    assert np.array_equal(pin["channels"].view(np.uint32), lens["channels"].view(np.uint32))
    assert np.array_equal(pin["rgba8"], lens["rgba8"])
    # every sample of a lens cell's pixels is visited, not just the bound's
    assert lens["stats"]["spl_count"] > pin["stats"]["spl_count"]


@pytest.mark.parametrize("z,focal", [(6.0, 12.0), (20.0, 8.0)])
def test_defocused_quad_matches_the_thin_lens(oracle, z, focal):
    o = _dof_opts(xsamples=4, ysamples=4)
    o.dof[:] = [1.4, 2.0, focal]
    rect = (17.3, 41.9, 11.6, 33.2)
    p = scenes.raster_quad(o, rect, z, (1.0, 0.5, 0.25))
    r = oracle.render(o, [p], occlusion_cull=True)
    ew, eh, spp = r["ext"]
    xy, _ = oracle.sample_positions(o)
    lens, cell = oracle.sample_lens(o)
    _, cells, _ = oracle.dof(o, np.array([z], np.float32))
    cx, cy = _coc(o, z)
    assert cx > 1.0                                 # visibly blurred
    q = xy.astype(np.float64) + lens.astype(np.float64) * np.array([cx, cy])
    d = np.minimum.reduce([q[..., 0] - rect[0], rect[1] - q[..., 0], q[..., 1] - rect[2], rect[3] - q[..., 1]])
    hit = (r["mp_id"] != 0xffffffff).reshape(eh, ew, spp)
    # A sample is offered to a micropolygon through its lens cell's bound of
    # offsets (CalculateDofBounds).  Those bounds come from the cells' corners
    # and do not hold every offset of the cell (projectToCircle bulges the
    # square's edges), so a sample whose offset lies outside its cell's bound
    # may go unvisited near the quad's edges: the lens law is checked on the
    # others.
    cb = cells[cell]
    inside_cell = np.all((lens >= cb[..., :2]) & (lens <= cb[..., 2:]), axis=-1)
    assert inside_cell.mean() > 0.6
    clear = (np.abs(d) > 1e-3) & inside_cell
    assert np.array_equal(hit[clear], (d > 0)[clear])
    # outside their cells' bounds, samples are hit only where the lens law says
    out = ~inside_cell & (np.abs(d) > 1e-3)
    assert not np.any(hit[out] & (d[out] < 0))
    dz = r["depth"][r["mp_id"] != 0xffffffff]
    assert dz.size > 1000 and np.allclose(dz, z, rtol=1e-5)


def test_dicing_rate_follows_the_coc_area(oracle):
    """A 12 x 12 pixel bilinear quad: 144 micropolygons in focus; out of
    focus the shading rate is multiplied by max(1, 0.025 * focusfactor *
    MinCoC^2), so each side is diced lround(12 / sqrt(rate)) times."""
    o = _dof_opts()
    z = 4.0
    p = scenes.raster_quad(o, (20.0, 32.0, 14.0, 26.0), z, (1.0, 1.0, 1.0))
    assert oracle.render(o, [p])["stats"]["micropolygons"] == 144
    o.dof[:] = [1.0, 3.0, 20.0]
    cmin = min(_coc(o, z))
    for ff in (0.0, 1.0, 0.5):
        p.focus_factor = ff
        rate = max(1.0, 0.025 * ff * cmin * cmin)
        side = max(int(math.floor(12.0 / math.sqrt(rate) + 0.5)), 1)
        assert rate == 1.0 or side < 12
        assert oracle.render(o, [p])["stats"]["micropolygons"] == side * side


def test_motion_keys_with_depth_of_field_are_refused(oracle):
    o = _dof_opts()
    o.dof[:] = [2.0, 1.0, 10.0]
    o.shutter_open, o.shutter_close = 0.0, 1.0
    p = scenes.raster_quad(o, (10.0, 20.0, 10.0, 20.0), 8.0, (1.0, 1.0, 1.0))
    p.nkeys = 2
    p.key_time[:] = [0.0, 1.0]
    for i in range(4):
        p.P1[i][:] = list(p.P[i])
    with pytest.raises(ValueError):
        oracle.render(o, [p])
