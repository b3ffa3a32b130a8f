"""ShadingInterpolation (CPU): the colour of a hit is the micropolygon's
first vertex colour under "constant" and the bilinear interpolation of its
four vertex colours at the hit's (u, v) under "smooth", the default
(CqMicroPolygon::InterpolateOutputs, micropolygon.cpp:1443-1462; attribute
default attributes.cpp:118).  The constant shader's colours are the same at
every vertex of a surface, so a constant surface stores exactly Os * Cs and a
smooth one ((w0 c + w1 c) + w2 c) + w3 c, which rounds differently from c at
some samples."""
import numpy as np
import pytest

from aqsis_amd import scenes
from aqsis_amd.abi import default_options


def _quad_scene(interp):
    o = default_options(xres=48, yres=40, fov=45.0, xsamples=4, ysamples=4, filter="box", filter_xwidth=1.0,
                        filter_ywidth=1.0)
    cs = (0.3, 0.7, 0.11)
    p = scenes.raster_quad(o, (5.3, 41.7, 4.1, 35.9), 9.0, cs)
    p.shading_interpolation = interp
    return o, [p], np.float32(1.0) * np.array(cs, np.float32)


def _hit_colours(oracle, o, prims):
    r = oracle.render(o, prims, occlusion_cull=True)
    hit = r["mp_id"] != 0xffffffff
    assert hit.sum() > 10000
    return r["rgbaz"][hit]


def test_constant_interpolation_stores_the_surface_colour(oracle):
    o, prims, c = _quad_scene(1)
    rgbaz = _hit_colours(oracle, o, prims)
    assert np.array_equal(rgbaz[:, 0:3], np.broadcast_to(c, (len(rgbaz), 3)))
    assert np.all(rgbaz[:, 3:6] == 1.0)


def test_smooth_interpolation_rounds_the_bilinear_sum(oracle):
    o, prims, c = _quad_scene(0)
    rgbaz = _hit_colours(oracle, o, prims)
    # within a few ulps of the surface colour, and not equal to it everywhere
    assert np.allclose(rgbaz[:, 0:3], c, rtol=4e-7, atol=0)
    assert not np.array_equal(rgbaz[:, 0:3], np.broadcast_to(c, (len(rgbaz), 3)))
    # the four weights sum to 1 up to rounding: (1-u)(1-v) + u(1-v) + (1-u)v + uv
    u = np.linspace(0.0, 1.0, 97, dtype=np.float32)
    U, V = np.meshgrid(u, u)
    one = np.float32(1)
    w = [(one - U) * (one - V), U * (one - V), (one - U) * V, U * V]
    for k in range(3):
        ck = c[k]
        s = ((w[0] * ck + w[1] * ck) + w[2] * ck) + w[3] * ck
        assert np.allclose(s, ck, rtol=4e-7, atol=0)


def test_default_is_smooth():
    p = scenes.raster_quad(default_options(xres=8, yres=8), (1, 6, 1, 6), 5.0, (1.0, 1.0, 1.0))
    assert p.shading_interpolation == 0


def test_unknown_interpolation_is_refused():
    from aqsis_amd.hider import HiderError, estimate
    o, prims, _ = _quad_scene(2)
    with pytest.raises(HiderError, match="shading_interpolation"):
        estimate(o, prims)
