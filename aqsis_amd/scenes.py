"""Seeded synthetic scenes for the benchmark configurations.

Every scene is a list of camera-space surfaces (RhPrim) plus the frame's
options; the same bytes feed the HIP hider and any CPU implementation of the
ABI.  The camera sits at the origin looking down +z (RenderMan camera
space) with Projection "perspective".  `to_rib` renders a scene as RIB text
for inspection.
"""
import ctypes as C
import math

import numpy as np

from .abi import (RhPrim, default_options, FLT_MAX, RH_PRIM_SPHERE, RH_PRIM_BICUBIC,
                  RH_PRIM_BILINEAR)


def _identity16():
    return [1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0]


def _translate16(x, y, z):
    m = _identity16()
    m[12], m[13], m[14] = float(x), float(y), float(z)
    return m


def prim_array(prims):
    arr = (RhPrim * len(prims))()
    for i, p in enumerate(prims):
        arr[i] = p
    return arr


def sphere(radius, centre, Cs, Os=(1.0, 1.0, 1.0), shading_rate=1.0):
    p = RhPrim()
    p.type = RH_PRIM_SPHERE
    p.nkeys = 1
    p.shading_rate = shading_rate
    p.focus_factor = 1.0
    p.Cs[:] = [float(c) for c in Cs]
    p.Os[:] = [float(c) for c in Os]
    p.sphere[:] = [float(radius), -float(radius), float(radius), 360.0]
    p.xform[0][:] = _translate16(*centre)
    p.xform[1][:] = _translate16(*centre)
    return p


def bicubic(P, Cs, Os=(1.0, 1.0, 1.0), shading_rate=1.0):
    p = RhPrim()
    p.type = RH_PRIM_BICUBIC
    p.nkeys = 1
    p.shading_rate = shading_rate
    p.focus_factor = 1.0
    p.Cs[:] = [float(c) for c in Cs]
    p.Os[:] = [float(c) for c in Os]
    P = np.asarray(P, dtype=np.float32).reshape(16, 3)
    for i in range(16):
        p.P[i][:] = [float(v) for v in P[i]]
    p.xform[0][:] = _identity16()
    p.xform[1][:] = _identity16()
    return p


def bilinear(P, Cs, Os=(1.0, 1.0, 1.0), shading_rate=1.0):
    p = RhPrim()
    p.type = RH_PRIM_BILINEAR
    p.nkeys = 1
    p.shading_rate = shading_rate
    p.focus_factor = 1.0
    p.Cs[:] = [float(c) for c in Cs]
    p.Os[:] = [float(c) for c in Os]
    P = np.asarray(P, dtype=np.float32).reshape(4, 3)
    for i in range(4):
        p.P[i][:] = [float(v) for v in P[i]]
    p.xform[0][:] = _identity16()
    p.xform[1][:] = _identity16()
    return p


def spheres_scene(seed=1, n=64, res=(256, 256), fov=45.0, depth=15.0,
                  samples=(4, 4), filt=("gaussian", 2.0, 2.0)):
    """Plain hider, CPU-runnable case: 64 spheres with centres uniform in a
    [-5,5]^3 box moved `depth` units in front of the camera, radii uniform in
    [0.3, 1.2], constant shader, per-object Cs uniform in [0,1]^3, Os = 1."""
    rng = np.random.default_rng(seed)
    opts = default_options(xres=res[0], yres=res[1], fov=fov,
                           xsamples=samples[0], ysamples=samples[1],
                           filter=filt[0], filter_xwidth=filt[1],
                           filter_ywidth=filt[2])
    prims = []
    for _ in range(n):
        c = rng.uniform(-5.0, 5.0, 3)
        c[2] += depth
        r = rng.uniform(0.3, 1.2)
        cs = rng.uniform(0.0, 1.0, 3)
        prims.append(sphere(np.float32(r), np.float32(c), np.float32(cs)))
    return opts, prims


def _bezier_patch(rng, centre, size, jitter):
    """A random planar 4x4 Bezier lattice of side `size` centred at `centre`,
    its normal biased toward the camera, plus N(0, jitter*size) noise on
    every control point."""
    n = rng.normal(size=3) + np.array([0.0, 0.0, -2.0])
    n /= np.linalg.norm(n)
    r = rng.normal(size=3)
    u_axis = np.cross(n, r)
    u_axis /= np.linalg.norm(u_axis)
    v_axis = np.cross(n, u_axis)
    P = np.empty((16, 3))
    for iv in range(4):
        for iu in range(4):
            s = (iu / 3.0 - 0.5) * size
            q = (iv / 3.0 - 0.5) * size
            P[iv * 4 + iu] = centre + s * u_axis + q * v_axis
    P += rng.normal(0.0, jitter, P.shape) * size
    return P.astype(np.float32)


def patches_scene(seed=1, n=10000, res=(1920, 1080), fov=45.0, znear=8.0,
                  zfar=40.0, size_px=28.0, jitter=0.1, samples=(4, 4),
                  filt=("catmull-rom", 3.0, 3.0), opacity=None, shading_rate=1.0):
    """Plain hider at HD: `n` bicubic Bezier patches placed uniformly in the
    view frustum between znear and zfar, each about `size_px` raster pixels
    across, with N(0, jitter) control-point noise (relative to the patch
    size) and a constant Cs per patch.  opacity=(lo, hi) draws Os uniform per
    patch; None gives opaque patches."""
    rng = np.random.default_rng(seed)
    opts = default_options(xres=res[0], yres=res[1], fov=fov,
                           xsamples=samples[0], ysamples=samples[1],
                           filter=filt[0], filter_xwidth=filt[1],
                           filter_ywidth=filt[2])
    aspect = res[0] / res[1]
    tanh = math.tan(math.radians(fov) / 2.0)
    # screen window [-aspect, aspect] x [-1, 1] for aspect >= 1
    sx = tanh * max(aspect, 1.0)
    sy = tanh * max(1.0 / aspect, 1.0)
    px_per_unit_at_1 = (res[0] / 2.0) / sx
    prims = []
    for _ in range(n):
        z = rng.uniform(znear, zfar)
        x = rng.uniform(-sx, sx) * z
        y = rng.uniform(-sy, sy) * z
        size = size_px * z / px_per_unit_at_1
        P = _bezier_patch(rng, np.array([x, y, z]), size, jitter)
        cs = rng.uniform(0.0, 1.0, 3)
        if opacity is None:
            os_ = (1.0, 1.0, 1.0)
        else:
            o = rng.uniform(opacity[0], opacity[1])
            os_ = (o, o, o)
        prims.append(bicubic(P, np.float32(cs), np.float32(os_), shading_rate))
    return opts, prims


def raster_quad(opts, rect, z, Cs, Os=(1.0, 1.0, 1.0), shading_rate=1.0):
    """A bilinear patch facing the camera at camera depth z whose raster
    projection is the rectangle rect = (x0, x1, y0, y1) of the frame `opts`
    describes (perspective, pixel aspect 1)."""
    W, H = opts.xres, opts.yres
    aspect = W / H
    tanh = math.tan(math.radians(opts.fov) / 2.0)
    sx = tanh * max(aspect, 1.0)
    sy = tanh * max(1.0 / aspect, 1.0)

    def cam(rx, ry):
        return [(rx / W * 2.0 - 1.0) * sx * z, (1.0 - ry / H * 2.0) * sy * z, z]

    x0, x1, y0, y1 = rect
    P = [cam(x0, y0), cam(x1, y0), cam(x0, y1), cam(x1, y1)]
    return bilinear(P, Cs, Os, shading_rate)


def coplanar_scene(res=(64, 32), samples=(2, 2), filt=("box", 1.0, 1.0)):
    """Exact depth ties that occlusion culling reorders: a near quad covering
    bucket 0, and two coplanar quads at z = 10 that overlap in bucket 1, the
    first starting in bucket 0 (behind the near quad there, so aqsis culls it
    and reposts it to bucket 1) and the second starting in bucket 1."""
    opts = default_options(xres=res[0], yres=res[1], fov=90.0, xsamples=samples[0], ysamples=samples[1],
                           filter=filt[0], filter_xwidth=filt[1], filter_ywidth=filt[2])
    prims = [raster_quad(opts, (-4.0, 18.0, -4.0, 18.0), 2.0, (0.9, 0.1, 0.1)),
             raster_quad(opts, (4.0, 28.0, 4.0, 12.0), 10.0, (0.1, 0.9, 0.1)),
             raster_quad(opts, (20.0, 40.0, 2.0, 14.0), 10.0, (0.1, 0.1, 0.9))]
    return opts, prims


def decal_scene(seed=1, n=60, res=(128, 96), samples=(2, 2), filt=("gaussian", 2.0, 2.0), planes=(6.0, 9.0, 12.0),
                opacity=None, occluders=6):
    """Coplanar geometry at many places: `n` camera-facing quads of random
    raster size, each on one of a few depth planes (so quads on one plane tie
    exactly wherever they overlap), behind `occluders` nearer quads that
    cover whole buckets, so occlusion culling reposts surfaces.  opacity=(lo,
    hi) makes the decals transparent (equal-depth transparent hits)."""
    rng = np.random.default_rng(seed)
    opts = default_options(xres=res[0], yres=res[1], fov=60.0, xsamples=samples[0], ysamples=samples[1],
                           filter=filt[0], filter_xwidth=filt[1], filter_ywidth=filt[2])
    prims = []
    for _ in range(occluders):
        bx = 16.0 * rng.integers(0, max(1, res[0] // 16))
        by = 16.0 * rng.integers(0, max(1, res[1] // 16))
        w = 16.0 * rng.integers(1, 3)
        prims.append(raster_quad(opts, (bx - 3.0, bx + w + 3.0, by - 3.0, by + 19.0), 3.0,
                                 np.float32(rng.uniform(0, 1, 3))))
    for _ in range(n):
        w, h = rng.uniform(6.0, 40.0, 2)
        x0 = rng.uniform(-10.0, res[0] - 5.0)
        y0 = rng.uniform(-10.0, res[1] - 5.0)
        z = float(planes[rng.integers(0, len(planes))])
        if opacity is None:
            os_ = (1.0, 1.0, 1.0)
        else:
            o = rng.uniform(opacity[0], opacity[1])
            os_ = (o, o, o)
        prims.append(raster_quad(opts, (x0, x0 + w, y0, y0 + h), z, np.float32(rng.uniform(0, 1, 3)),
                                 np.float32(os_)))
    return opts, prims


def eye_plane_scene(seed=1, res=(96, 64), samples=(2, 2), filt=("gaussian", 2.0, 2.0), n=40):
    """Surfaces that cross the eye plane: two ground planes (a bicubic and a
    bilinear patch) running from behind the camera (z = -6) to z = 30 below
    the view axis, which CullSurface marks undiceable and splits until the
    pieces lie in front of the camera or behind it, or exceed the eye-split
    limit (imagebuffer.cpp:284-300); plus `n` ordinary patches."""
    rng = np.random.default_rng(seed)
    opts, prims = patches_scene(seed=seed, n=n, res=res, samples=samples, filt=filt, znear=6.0, zfar=20.0)
    xs = np.linspace(-12.0, 12.0, 4, dtype=np.float32)
    zs = np.linspace(-6.0, 30.0, 4, dtype=np.float32)
    P = np.array([[x, -2.5 + 0.3 * np.sin(x + z), z] for z in zs for x in xs], np.float32)
    prims.append(bicubic(P, np.float32(rng.uniform(0, 1, 3))))
    Q = np.array([[-8.0, 2.0, -4.0], [8.0, 2.0, -4.0], [-8.0, 1.5, 25.0], [8.0, 1.5, 25.0]], np.float32)
    prims.append(bilinear(Q, np.float32(rng.uniform(0, 1, 3))))
    return opts, prims


def motion_scene(seed=1, n=600, res=(256, 256), fov=45.0, znear=8.0, zfar=40.0, size_px=28.0,
                 jitter=0.1, samples=(8, 8), filt=("gaussian", 2.0, 2.0), opacity=(0.2, 0.8),
                 shift_sigma=0.05, shutter=(0.0, 1.0), key_times=(0.0, 1.0), shading_rate=1.0):
    """Transparency and motion blur: `n` bicubic patches as patches_scene,
    each a deforming surface whose second key (RiMotionBegin [t0 t1] with the
    patch twice) is the first moved in the image plane by N(0, shift_sigma *
    width) raster pixels per axis, shutter open over the keys.  Os uniform in
    `opacity`."""
    rng = np.random.default_rng(seed)
    opts, prims = patches_scene(seed=seed, n=n, res=res, fov=fov, znear=znear, zfar=zfar, size_px=size_px,
                                jitter=jitter, samples=samples, filt=filt, opacity=opacity,
                                shading_rate=shading_rate)
    opts.shutter_open, opts.shutter_close = float(shutter[0]), float(shutter[1])
    aspect = res[0] / res[1]
    sx = math.tan(math.radians(fov) / 2.0) * max(aspect, 1.0)
    px_per_unit_at_1 = (res[0] / 2.0) / sx
    for p in prims:
        P = np.array([list(p.P[i]) for i in range(16)], dtype=np.float64)
        z = float(P[:, 2].mean())
        d = rng.normal(0.0, shift_sigma * res[0], 2) * z / px_per_unit_at_1   # raster px -> camera units
        P1 = P + np.array([d[0], -d[1], 0.0])
        p.nkeys = 2
        p.key_time[:] = [float(key_times[0]), float(key_times[1])]
        for i in range(16):
            p.P1[i][:] = [float(v) for v in np.float32(P1[i])]
    return opts, prims


def _prim_records(n):
    """A ctypes array of n RhPrim, zeroed but for focus_factor (1), and a numpy
    view of its bytes."""
    arr = (RhPrim * n)()
    raw = np.frombuffer(arr, dtype=np.uint8).reshape(n, C.sizeof(RhPrim))
    _field(raw, "focus_factor", ())[:] = 1.0   # GeometricApproximation "focusfactor" default
    return arr, raw


def _field(raw, name, shape, dtype=np.float32):
    off = getattr(RhPrim, name).offset
    size = int(np.prod(shape)) * np.dtype(dtype).itemsize
    return raw[:, off:off + size].view(dtype).reshape((raw.shape[0],) + tuple(shape))


def bicubic_records(P, Cs, Os, shading_rate=1.0):
    """Bicubic patches as a ctypes RhPrim array, built without a Python loop
    per patch: P (n, 16, 3), Cs (n, 3), Os (n, 3)."""
    P = np.asarray(P, np.float32)
    n = P.shape[0]
    arr, raw = _prim_records(n)
    _field(raw, "type", (), np.int32)[:] = RH_PRIM_BICUBIC
    _field(raw, "nkeys", (), np.int32)[:] = 1
    _field(raw, "shading_rate", ())[:] = shading_rate
    _field(raw, "Cs", (3,))[:] = np.asarray(Cs, np.float32)
    _field(raw, "Os", (3,))[:] = np.asarray(Os, np.float32)
    eye = np.asarray(_identity16(), np.float32)
    _field(raw, "xform", (2, 16))[:] = eye
    _field(raw, "P", (16, 3))[:] = P
    return arr


def clustered_scene(seed=1, res=(2048, 2048), bucket=16, median=2000.0, sigma=1.5, cap=200000,
                    patch_px=16.0, fov=45.0, znear=8.0, zfar=40.0, jitter=0.1, samples=(4, 4),
                    filt=("gaussian", 2.0, 2.0), opacity=None):
    """The batched configuration: bicubic patches clustered so that the
    micropolygon count of each bucket follows LogNormal(ln median, sigma),
    clipped at `cap`.  A patch about `patch_px` pixels across dices to about
    patch_px^2 micropolygons at ShadingRate 1, so bucket b receives
    round(N_b / patch_px^2) patches (N_b its drawn count; a fractional part
    is kept with that probability), centred in the bucket with N(0, bucket/4)
    pixel offsets, at uniform depths in [znear, zfar].  Returns (opts, prims)
    with prims a ctypes RhPrim array."""
    rng = np.random.default_rng(seed)
    opts = default_options(xres=res[0], yres=res[1], fov=fov, xsamples=samples[0], ysamples=samples[1],
                           filter=filt[0], filter_xwidth=filt[1], filter_ywidth=filt[2],
                           bucket_x=bucket, bucket_y=bucket)
    nbx, nby = (res[0] + bucket - 1) // bucket, (res[1] + bucket - 1) // bucket
    counts = np.minimum(rng.lognormal(np.log(median), sigma, nbx * nby), cap)
    per = counts / (patch_px * patch_px)
    npatch = np.floor(per).astype(np.int64)
    npatch += rng.uniform(size=per.shape) < (per - npatch)
    bidx = np.repeat(np.arange(nbx * nby), npatch)
    n = bidx.size
    cx = (bidx % nbx) * bucket + bucket / 2.0 + rng.normal(0.0, bucket / 4.0, n)
    cy = (bidx // nbx) * bucket + bucket / 2.0 + rng.normal(0.0, bucket / 4.0, n)
    return opts, _patch_records(rng, cx, cy, res, fov, znear, zfar, patch_px, jitter, opacity)


def _patch_records(rng, cx, cy, res, fov, znear, zfar, patch_px, jitter, opacity):
    """Bicubic patches about `patch_px` pixels across centred on raster
    points (cx, cy) at uniform depths in [znear, zfar]: a random planar 4x4
    lattice with its normal biased toward the camera plus N(0, jitter) noise
    relative to its size, Cs uniform, Os = 1 or uniform in `opacity`."""
    n = cx.size
    aspect = res[0] / res[1]
    tanh = np.tan(np.radians(fov) / 2.0)
    sx = tanh * max(aspect, 1.0)
    sy = tanh * max(1.0 / aspect, 1.0)
    z = rng.uniform(znear, zfar, n)
    x = (cx / res[0] * 2.0 - 1.0) * sx * z
    y = (1.0 - cy / res[1] * 2.0) * sy * z
    px_per_unit_at_1 = (res[0] / 2.0) / sx
    size = patch_px * z / px_per_unit_at_1
    nrm = rng.normal(size=(n, 3)) + np.array([0.0, 0.0, -2.0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    r = rng.normal(size=(n, 3))
    ua = np.cross(nrm, r)
    ua /= np.linalg.norm(ua, axis=1, keepdims=True)
    va = np.cross(nrm, ua)
    g = np.arange(4) / 3.0 - 0.5
    su = np.tile(g, 4)             # iu varies fastest
    sv = np.repeat(g, 4)
    centre = np.stack([x, y, z], axis=1)
    P = (centre[:, None, :] + (su[None, :, None] * size[:, None, None]) * ua[:, None, :]
         + (sv[None, :, None] * size[:, None, None]) * va[:, None, :])
    P += rng.normal(0.0, jitter, P.shape) * size[:, None, None]
    cs = rng.uniform(0.0, 1.0, (n, 3))
    if opacity is None:
        os_ = np.ones((n, 3))
    else:
        o = rng.uniform(opacity[0], opacity[1], n)
        os_ = np.repeat(o[:, None], 3, axis=1)
    return bicubic_records(P, cs, os_)


def serving_requests(seed=1, count=2000, sides=(256, 512, 1024, 2048), spp=(2, 4, 8), patches=(100, 20000),
                     opacity=(0.2, 1.0), rate_hz=None):
    """The serving stream: `count` render requests, each a square frame of a
    side drawn from `sides`, PixelSamples n x n with n from `spp`, a number
    of bicubic patches uniform in `patches` (about 28 pixels across at
    1920 pixels of width, scaled with the side), Os uniform in `opacity`,
    gaussian 2x2.  With rate_hz, arrival times of a Poisson process of that
    rate are returned as well.  Yields (opts, prims, arrival_s)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    for _ in range(count):
        side = int(rng.choice(sides))
        n_s = int(rng.choice(spp))
        npatch = int(rng.integers(patches[0], patches[1] + 1))
        opts = default_options(xres=side, yres=side, fov=45.0, xsamples=n_s, ysamples=n_s, filter="gaussian",
                               filter_xwidth=2.0, filter_ywidth=2.0)
        cx = rng.uniform(0.0, side, npatch)
        cy = rng.uniform(0.0, side, npatch)
        prims = _patch_records(rng, cx, cy, (side, side), 45.0, 8.0, 40.0, 28.0 * side / 1920.0, 0.1, opacity)
        if rate_hz:
            t += rng.exponential(1.0 / rate_hz)
        yield opts, prims, t


_DEPTH_NAMES = ["min", "midpoint", "max", "average"]


def to_rib(opts, prims):
    """RIB text of a scene in camera space (spheres as a translation), which
    aqsis_amd.rib.parse reads back to the same options and records; two-key
    surfaces as a MotionBegin block of two patches or two translations."""
    names = {v: k for k, v in __import__("aqsis_amd.abi", fromlist=["FILTERS"]).FILTERS.items()}
    out = [f'Format {opts.xres} {opts.yres} {opts.pixel_aspect}',
           f'PixelSamples {opts.xsamples} {opts.ysamples}',
           f'PixelFilter "{names[opts.filter]}" {opts.filter_xwidth} {opts.filter_ywidth}',
           f'Exposure {opts.exposure_gain} {opts.exposure_gamma}',
           f'Quantize "rgba" {opts.quantize_one} {opts.quantize_min} {opts.quantize_max} {opts.quantize_dither}',
           f'CropWindow {opts.crop[0]} {opts.crop[1]} {opts.crop[2]} {opts.crop[3]}',
           f'Clipping {opts.clip_near} {opts.clip_far}',
           f'Option "limits" "bucketsize" [{opts.bucket_x} {opts.bucket_y}] "eyesplits" [{opts.eyesplits}]',
           f'Option "limits" "zthreshold" [{opts.zthreshold[0]} {opts.zthreshold[1]} {opts.zthreshold[2]}]',
           f'Hider "hidden" "jitter" [{opts.jitter}] "depthfilter" "{_DEPTH_NAMES[opts.depth_filter]}"',
           f'Shutter {opts.shutter_open} {opts.shutter_close}',
           (f'DepthOfField {opts.dof[0]} {opts.dof[1]} {opts.dof[2]}' if opts.dof[0] < FLT_MAX else 'DepthOfField'),
           f'Projection "perspective" "fov" [{opts.fov}]',
           'WorldBegin', 'Surface "constant"']
    # display requests, unless the default single one without z: one per
    # num_displays, z in the first one's mode with display_z
    if opts.display_z or opts.num_displays != 1:
        displays = [f'Display "frame" "null" "{"rgbaz" if opts.display_z else "rgba"}"']
        displays += [f'Display "+frame{i}" "null" "rgba"' for i in range(1, opts.num_displays)]
        out[0:0] = displays
    for p in prims:
        out.append('AttributeBegin')
        out.append(f'Color [{p.Cs[0]} {p.Cs[1]} {p.Cs[2]}]')
        out.append(f'Opacity [{p.Os[0]} {p.Os[1]} {p.Os[2]}]')
        out.append(f'ShadingRate {p.shading_rate}')
        out.append(f'GeometricApproximation "focusfactor" {p.focus_factor}')
        out.append('ShadingInterpolation "constant"' if p.shading_interpolation else 'ShadingInterpolation "smooth"')
        out.append(f'Attribute "cull" "hidden" [{0 if p.no_cull_hidden else 1}]')
        out.append(f'Matte {p.matte}')
        moving = p.nkeys == 2
        if moving:
            motion = f'MotionBegin [{p.key_time[0]} {p.key_time[1]}]'
        if p.type == RH_PRIM_SPHERE:
            keys = [list(p.xform[k]) for k in range(2 if moving else 1)]
            if moving:
                out.append(motion)
            for m in keys:
                out.append(f'Translate {m[12]} {m[13]} {m[14]}')
            if moving:
                out.append('MotionEnd')
            out.append(f'Sphere {p.sphere[0]} {p.sphere[1]} {p.sphere[2]} {p.sphere[3]}')
        else:
            n = 16 if p.type == RH_PRIM_BICUBIC else 4
            kind = "bicubic" if p.type == RH_PRIM_BICUBIC else "bilinear"
            if moving:
                out.append(motion)
            for P in ([p.P, p.P1] if moving else [p.P]):
                pts = " ".join(f"{P[i][0]} {P[i][1]} {P[i][2]}" for i in range(n))
                out.append(f'Patch "{kind}" "P" [{pts}]')
            if moving:
                out.append('MotionEnd')
        out.append('AttributeEnd')
    out.append('WorldEnd')
    return "\n".join(out) + "\n"
