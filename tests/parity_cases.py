"""Scenes used by the parity tests: small enough for the oracle to finish in
seconds, covering every option the HIP path implements."""
import numpy as np

from aqsis_amd import scenes
from aqsis_amd.abi import default_options


def spheres_config():
    # The CPU-runnable configuration in full: 256x256, 16x16 buckets,
    # PixelSamples 4x4, gaussian 2x2, ShadingRate 1, 64 spheres.
    return scenes.spheres_scene(seed=1)


def patches_small():
    # The HD configuration's scene generator, filter and sampling at a
    # quarter of the linear size.
    return scenes.patches_scene(seed=7, n=700, res=(480, 270))


def patches_8x8():
    return scenes.patches_scene(seed=3, n=150, res=(160, 96), samples=(8, 8), filt=("gaussian", 2.0, 2.0))


def spheres_box_2x2():
    return scenes.spheres_scene(seed=5, n=40, res=(200, 150), samples=(2, 2), filt=("box", 1.0, 1.0))


def patches_crop_odd():
    opts, prims = scenes.patches_scene(seed=11, n=500, res=(250, 130), filt=("triangle", 2.0, 2.0))
    opts.crop[:] = [0.13, 0.87, 0.21, 0.93]
    return opts, prims


def patches_gridsampler():
    opts, prims = scenes.patches_scene(seed=13, n=300, res=(192, 128), samples=(3, 3),
                                       filt=("sinc", 3.0, 3.0))
    opts.jitter = 0
    return opts, prims


def spheres_exposure_mitchell():
    opts, prims = scenes.spheres_scene(seed=17, n=48, res=(176, 144), samples=(4, 4),
                                       filt=("mitchell", 3.0, 3.0))
    opts.exposure_gain = 1.5
    opts.bucket_x, opts.bucket_y = 32, 16
    return opts, prims


def spheres_disk_clip_aspect():
    # Disk filter, PixelSamples 3x2, non-square pixels, clipping planes that
    # cut through the spheres' depth range (surfaces wholly outside are
    # culled, the rest are sampled), and two display requests drawing dither
    # values from the shared random stream.
    opts, prims = scenes.spheres_scene(seed=29, n=56, res=(168, 120), samples=(3, 2), filt=("disk", 2.0, 2.0))
    opts.pixel_aspect = 1.25
    opts.clip_near, opts.clip_far = 11.5, 17.0
    opts.num_displays = 2
    return opts, prims


def patches_bessel_1x1():
    # Bessel 3x3 with one sample per pixel, exposure gamma, no dither.
    opts, prims = scenes.patches_scene(seed=31, n=300, res=(176, 112), samples=(1, 1), filt=("bessel", 3.0, 3.0))
    opts.exposure_gamma = 2.2
    opts.quantize_dither = 0.0
    return opts, prims


def spheres_quantize_range():
    # Quantize "rgba" 200 10 230 0.5: a different one value and a clamp
    # range inside the byte.
    opts, prims = scenes.spheres_scene(seed=37, n=48, res=(128, 96), samples=(2, 2))
    opts.quantize_one, opts.quantize_min, opts.quantize_max, opts.quantize_dither = 200.0, 10.0, 230.0, 0.5
    return opts, prims


def spheres_wide_gaussian():
    # A 5x5 filter footprint (DiscreteShift 2) with 3x3 samples and a frame
    # that is not a whole number of 8x8 pixel tiles.
    return scenes.spheres_scene(seed=19, n=40, res=(150, 98), samples=(3, 3), filt=("gaussian", 4.0, 4.0))


def patches_transparent():
    # Os uniform in [0.2, 0.8] per patch: per-sample hit lists, depth sort and
    # back-to-front compositing (imagepixel.cpp:144-332).
    return scenes.patches_scene(seed=23, n=300, res=(192, 128), opacity=(0.2, 0.8),
                                filt=("gaussian", 2.0, 2.0))


def patches_transparent_8x8_zthreshold():
    # The transparent configuration's sampling (8x8, gaussian 2x2) on a small
    # frame, with zthreshold 0.5 so transparent hits also set the depth.
    opts, prims = scenes.patches_scene(seed=29, n=120, res=(96, 64), samples=(8, 8), opacity=(0.2, 0.8),
                                       filt=("gaussian", 2.0, 2.0))
    opts.zthreshold[:] = [0.5, 0.5, 0.5]
    return opts, prims


def spheres_mixed_opacity():
    opts, prims = scenes.spheres_scene(seed=31, n=48, res=(160, 120))
    for i, p in enumerate(prims):
        if i % 2:
            p.Os[:] = [0.5, 0.25, 0.75]
    return opts, prims


def patches_deep_transparent():
    # About six transparent layers per sample: more hits than the first
    # arena holds, so the frame is redone with a larger one.
    return scenes.patches_scene(seed=37, n=100, res=(128, 96), opacity=(0.2, 0.8), filt=("box", 1.0, 1.0))


def bilinear_patches():
    # Bilinear patches (CqSurfacePatchBilinear, BilinearEvaluate dicing):
    # corners of each of the HD generator's patches, mixed opacity.
    opts, prims = scenes.patches_scene(seed=43, n=250, res=(160, 120), filt=("catmull-rom", 3.0, 3.0))
    out = []
    for i, p in enumerate(prims):
        P = np.array([list(p.P[k]) for k in (0, 3, 12, 15)], np.float32)
        out.append(scenes.bilinear(P, list(p.Cs), (1.0, 1.0, 1.0) if i % 3 else (0.6, 0.6, 0.6), 1.0))
    return opts, out


def motion_transparent_8x8():
    # The transparency + motion blur configuration at half the linear size:
    # 8x8 samples, gaussian 2x2, Os 0.2-0.8, every patch a deforming surface
    # moved N(0, 0.05 W) pixels between keys 0 and 1, shutter [0 1].
    return scenes.motion_scene(seed=21, n=160, res=(128, 96))


def motion_mixed_opaque():
    # Moving and static opaque patches in one frame (4x4, catmull-rom 3x3),
    # keys at 0.25 and 0.75 inside a [0, 1] shutter, so sample times before
    # and after the keys use the end keys' vertices exactly.
    opts, prims = scenes.motion_scene(seed=22, n=220, res=(160, 112), samples=(4, 4), opacity=None,
                                      filt=("catmull-rom", 3.0, 3.0), shift_sigma=0.08, key_times=(0.25, 0.75))
    for i, p in enumerate(prims):
        if i % 3 == 0:
            p.nkeys = 1
    return opts, prims


def motion_spheres():
    # Moving spheres (key 1's object->camera matrix translated), 3x3 samples.
    opts, prims = scenes.spheres_scene(seed=23, n=40, res=(144, 112), samples=(3, 3), filt=("box", 1.0, 1.0))
    opts.shutter_open, opts.shutter_close = 0.0, 1.0
    rng = np.random.default_rng(23)
    for p in prims:
        d = rng.normal(0.0, 0.4, 3)
        d[2] = 0.0
        p.nkeys = 2
        p.key_time[:] = [0.0, 1.0]
        m = list(p.xform[0])
        m[12] += float(d[0]); m[13] += float(d[1])
        p.xform[1][:] = [float(np.float32(v)) for v in m]
    return opts, prims


def motion_closed_shutter():
    # Shutter open == close (the fast-shutter path: every sub-bound visits all
    # of a pixel's samples, all at one time between the keys, so a
    # transparent micropolygon is stored once per sub-bound holding the
    # sample), 2x2 samples; mostly opaque so that no sample holds more
    # transparent hits than the device keeps.
    opts, prims = scenes.motion_scene(seed=31, n=90, res=(96, 80), samples=(2, 2), filt=("gaussian", 2.0, 2.0),
                                      opacity=(0.5, 1.0), shutter=(0.4, 0.4), key_times=(0.0, 1.0),
                                      shift_sigma=0.06)
    for i, p in enumerate(prims):
        if i % 4:
            p.Os[:] = [1.0, 1.0, 1.0]
    return opts, prims


def motion_one_sample():
    # One sample per pixel: four sub-bounds at most (max(4, ns)), whose index
    # windows all hold the one sample; transparent and opaque patches mixed.
    opts, prims = scenes.motion_scene(seed=33, n=150, res=(112, 96), samples=(1, 1), filt=("box", 1.0, 1.0),
                                      opacity=(0.4, 1.0), shift_sigma=0.08)
    for p in prims[::2]:
        p.Os[:] = [1.0, 1.0, 1.0]
    return opts, prims


def coplanar_ties():
    # Exact depth ties that occlusion culling reorders (RepostSurface): the
    # culled quad is diced after the other and loses the ties.
    return scenes.coplanar_scene()


def decal_ties():
    # Many camera-facing quads on three depth planes (equal depths wherever
    # two overlap) behind near quads that cover whole buckets.
    return scenes.decal_scene(seed=1)


def decal_ties_transparent():
    # The same with transparent quads: equal-depth transparent hits, whose
    # order into the depth sort follows the dicing order.
    return scenes.decal_scene(seed=2, opacity=(0.3, 0.7), n=50)


def decal_ties_motion():
    # Ties and culling in a frame with moving surfaces (the motion sampler's
    # probes): the decal quads as two-key deforming surfaces, half of them
    # moved in the image plane.
    opts, prims = scenes.decal_scene(seed=3, n=45, samples=(2, 2))
    opts.shutter_open, opts.shutter_close = 0.0, 1.0
    rng = np.random.default_rng(3)
    for i, p in enumerate(prims):
        p.nkeys = 2
        p.key_time[:] = [0.0, 1.0]
        d = rng.normal(0.0, 0.3, 2) if i % 2 else np.zeros(2)
        for k in range(4):
            p.P1[k][:] = [float(np.float32(p.P[k][0] + d[0])), float(np.float32(p.P[k][1] + d[1])), p.P[k][2]]
    return opts, prims


def patches_cull_flag_mixed():
    # Attribute "cull" "hidden" [0] on every other surface: only the others
    # may be occlusion culled.
    opts, prims = scenes.patches_scene(seed=47, n=400, res=(240, 136))
    for p in prims[::2]:
        p.no_cull_hidden = 1
    return opts, prims


def eye_plane_split():
    # Ground planes crossing the eye plane (eye splits, pieces behind the
    # camera culled) under ordinary patches.
    return scenes.eye_plane_scene(seed=5)


def patches_dof():
    # Depth of field on the HD scene generator: the focal plane inside the
    # patches' depth range (some in focus, most blurred by up to several
    # pixels), the lens-cell sampler and the CoC-grown bounds and dicing rate.
    opts, prims = scenes.patches_scene(seed=53, n=260, res=(160, 120), samples=(3, 3),
                                       filt=("gaussian", 2.0, 2.0))
    opts.dof[:] = [2.8, 1.2, 16.0]
    return opts, prims


def spheres_dof_transparent():
    # Strong blur (fstop 1.4) of spheres of mixed opacity: transparent hits
    # under the lens, a catmull-rom filter and 4x4 samples.
    opts, prims = scenes.spheres_scene(seed=59, n=40, res=(144, 112), samples=(4, 4),
                                       filt=("catmull-rom", 3.0, 3.0))
    rng = np.random.default_rng(59)
    for p in prims[1::2]:
        o = float(rng.uniform(0.3, 0.8))
        p.Os[:] = [o, o, o]
    opts.dof[:] = [1.4, 1.0, 12.0]
    return opts, prims


def decal_dof():
    # Equal-depth ties and occlusion culling under the lens: the decal quads
    # with the focal plane on the middle plane (in focus there, blurred on
    # the others), focus factor 0.5 on half of them.
    opts, prims = scenes.decal_scene(seed=4, n=50)
    opts.dof[:] = [2.0, 1.0, 9.0]
    for p in prims[::2]:
        p.focus_factor = 0.5
    return opts, prims


def patches_dof_1x1_crop():
    # One sample per pixel (one lens cell) with a crop window and a box filter.
    opts, prims = scenes.patches_scene(seed=61, n=300, res=(180, 120), samples=(1, 1), filt=("box", 1.0, 1.0))
    opts.crop[:] = [0.1, 0.9, 0.15, 0.85]
    opts.dof[:] = [2.0, 1.5, 20.0]
    return opts, prims


def deep_stack_100():
    # 100 transparent camera-facing quads over the middle of the frame, some
    # at equal depths: samples with up to 100 transparent hits (more than
    # the combine kernel sorts in LDS: the long-list pass, std::sort's full
    # introsort with ties).
    opts = default_options(xres=64, yres=48, fov=45.0, xsamples=2, ysamples=2, filter="box",
                           filter_xwidth=1.0, filter_ywidth=1.0)
    rng = np.random.default_rng(71)
    prims = []
    for i in range(100):
        x0, y0 = rng.uniform(4.0, 14.0), rng.uniform(4.0, 12.0)
        x1, y1 = rng.uniform(48.0, 60.0), rng.uniform(34.0, 44.0)
        z = 6.0 + 0.05 * (i - i % 3)            # groups of three at one depth
        o = float(rng.uniform(0.02, 0.1))
        prims.append(scenes.raster_quad(opts, (x0, x1, y0, y1), z, tuple(rng.uniform(0.0, 1.0, 3)), (o, o, o)))
    return opts, prims


def constant_interpolation_mixed():
    # ShadingInterpolation "constant" on every third surface, smooth (the
    # default) on the others; transparent and opaque, static and moving,
    # with and without depth of field: colours taken as they are on the
    # constant surfaces, bilinearly interpolated from equal corners on the
    # smooth ones.
    opts, prims = scenes.patches_scene(seed=83, n=240, res=(160, 96), samples=(3, 3),
                                       opacity=(0.3, 1.0))
    for p in prims[::3]:
        p.shading_interpolation = 1
    return opts, prims


def constant_interpolation_motion():
    opts, prims = motion_mixed_opaque()
    for p in prims[1::2]:
        p.shading_interpolation = 1
    return opts, prims


def constant_interpolation_dof():
    opts, prims = spheres_dof_transparent()
    for p in prims[::2]:
        p.shading_interpolation = 1
    return opts, prims


def _z_display(opts, filt):
    from aqsis_amd.abi import DEPTH_FILTERS
    opts.display_z = 1
    opts.depth_filter = DEPTH_FILTERS[filt]
    return opts


def depth_midpoint_z():
    # A z display with the midpoint depth filter: the occluding hit is the
    # nearest opaque one and the occluding depth the second nearest
    # (StoreSample's special case); samples without a list average the two,
    # listed ones take the mean of the two nearest opaque depths.  Mixed
    # opacity, occlusion culling against the second nearest depth.
    opts, prims = scenes.patches_scene(seed=89, n=260, res=(160, 96), samples=(3, 3), opacity=(0.3, 1.0),
                                       filt=("box", 1.0, 1.0))
    for p in prims[1::2]:
        p.Os[:] = [1.0, 1.0, 1.0]
    return _z_display(opts, "midpoint"), prims


def depth_max_z():
    # A z display with the max depth filter: no hit is occlusion cullable,
    # every hit (hidden ones too) is listed and composited, the depth is the
    # farthest one with opacity at zthreshold.
    opts, prims = scenes.patches_scene(seed=97, n=160, res=(128, 80), samples=(2, 2), opacity=(0.4, 1.0),
                                       filt=("gaussian", 2.0, 2.0))
    for p in prims[::2]:
        p.Os[:] = [1.0, 1.0, 1.0]
    return _z_display(opts, "max"), prims


def depth_average_z():
    # The average depth filter (every hit listed): the mean depth of the hits
    # with ANY opacity channel at zthreshold (0.6), while the composite's
    # opaque depths need all three.  Every surface has its red opacity above
    # the threshold, so no sample averages over nothing.
    opts, prims = scenes.patches_scene(seed=101, n=160, res=(128, 80), samples=(2, 2), filt=("box", 1.0, 1.0))
    rng = np.random.default_rng(101)
    for p in prims:
        p.Os[:] = [float(np.float32(rng.uniform(0.65, 1.0))), float(np.float32(rng.uniform(0.2, 1.0))),
                   float(np.float32(rng.uniform(0.2, 1.0)))]
    opts.zthreshold[:] = [0.6, 0.6, 0.6]
    return _z_display(opts, "average"), prims


def depth_midpoint_no_z():
    # The midpoint filter without a z display: hits stored as usual, the
    # combine's midpoint of the two nearest opaque depths for listed samples.
    opts, prims = scenes.patches_scene(seed=103, n=200, res=(144, 96), samples=(2, 2), opacity=(0.3, 1.0))
    for p in prims[::3]:
        p.Os[:] = [1.0, 1.0, 1.0]
    opts.depth_filter = 1
    return opts, prims


def matte_mixed():
    # RiMatte: matte surfaces (1) cut out what lies behind them, opaque ones
    # to transparent black, transparent ones through the combine's lerps;
    # matte alpha surfaces (2) are stored as occluding hits whatever their
    # opacity.  Opaque, transparent, matte and matte alpha patches mixed.
    opts, prims = scenes.patches_scene(seed=107, n=260, res=(160, 96), samples=(3, 3), opacity=(0.3, 1.0))
    for i, p in enumerate(prims):
        if i % 2 == 0:
            p.Os[:] = [1.0, 1.0, 1.0]
        if i % 5 == 1:
            p.matte = 1
        elif i % 5 == 3:
            p.matte = 2
    return opts, prims


def matte_max_z_motion():
    # Matte surfaces with every hit listed (a z display with the max
    # filter), in a frame with moving surfaces.
    opts, prims = motion_mixed_opaque()
    rng = np.random.default_rng(109)
    for i, p in enumerate(prims):
        if i % 4 == 1:
            p.matte = 1
        elif i % 4 == 2:
            o = float(np.float32(rng.uniform(0.3, 0.8)))
            p.Os[:] = [o, o, o]
        elif i % 4 == 3:
            p.matte = 2
    return _z_display(opts, "max"), prims


def matte_midpoint_z_motion():
    # The midpoint store and matte surfaces in the motion sampler.
    opts, prims = motion_transparent_8x8()
    for i, p in enumerate(prims):
        if i % 3 == 0:
            p.Os[:] = [1.0, 1.0, 1.0]
        if i % 4 == 1:
            p.matte = 1
        elif i % 4 == 3:
            p.matte = 2
    return _z_display(opts, "midpoint"), prims


def matte_midpoint_z_dof():
    # The midpoint store and matte surfaces under the lens.
    opts, prims = spheres_dof_transparent()
    for i, p in enumerate(prims):
        if i % 3 == 0:
            p.matte = 1
        elif i % 3 == 1:
            p.matte = 2
    return _z_display(opts, "midpoint"), prims


def depth_average_z_dof():
    # Every hit listed under the lens (the average filter).
    opts, prims = patches_dof()
    rng = np.random.default_rng(113)
    for p in prims[::2]:
        p.Os[:] = [float(np.float32(rng.uniform(0.65, 1.0))), 0.3, 0.5]
    opts.zthreshold[:] = [0.6, 0.6, 0.6]
    return _z_display(opts, "average"), prims


CASES = {
    "spheres_config": spheres_config,
    "patches_small": patches_small,
    "patches_8x8": patches_8x8,
    "spheres_box_2x2": spheres_box_2x2,
    "patches_crop_odd": patches_crop_odd,
    "patches_gridsampler": patches_gridsampler,
    "spheres_exposure_mitchell": spheres_exposure_mitchell,
    "spheres_wide_gaussian": spheres_wide_gaussian,
    "spheres_disk_clip_aspect": spheres_disk_clip_aspect,
    "patches_bessel_1x1": patches_bessel_1x1,
    "spheres_quantize_range": spheres_quantize_range,
    "patches_transparent": patches_transparent,
    "patches_transparent_8x8_zthreshold": patches_transparent_8x8_zthreshold,
    "spheres_mixed_opacity": spheres_mixed_opacity,
    "patches_deep_transparent": patches_deep_transparent,
    "bilinear_patches": bilinear_patches,
    "motion_transparent_8x8": motion_transparent_8x8,
    "motion_mixed_opaque": motion_mixed_opaque,
    "motion_spheres": motion_spheres,
    "motion_closed_shutter": motion_closed_shutter,
    "motion_one_sample": motion_one_sample,
    "coplanar_ties": coplanar_ties,
    "decal_ties": decal_ties,
    "decal_ties_transparent": decal_ties_transparent,
    "decal_ties_motion": decal_ties_motion,
    "patches_cull_flag_mixed": patches_cull_flag_mixed,
    "eye_plane_split": eye_plane_split,
    "patches_dof": patches_dof,
    "spheres_dof_transparent": spheres_dof_transparent,
    "decal_dof": decal_dof,
    "patches_dof_1x1_crop": patches_dof_1x1_crop,
    "deep_stack_100": deep_stack_100,
    "constant_interpolation_mixed": constant_interpolation_mixed,
    "constant_interpolation_motion": constant_interpolation_motion,
    "constant_interpolation_dof": constant_interpolation_dof,
    "depth_midpoint_z": depth_midpoint_z,
    "depth_max_z": depth_max_z,
    "depth_average_z": depth_average_z,
    "depth_midpoint_no_z": depth_midpoint_no_z,
    "matte_mixed": matte_mixed,
    "matte_max_z_motion": matte_max_z_motion,
    "matte_midpoint_z_motion": matte_midpoint_z_motion,
    "matte_midpoint_z_dof": matte_midpoint_z_dof,
    "depth_average_z_dof": depth_average_z_dof,
}
