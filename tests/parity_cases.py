"""Scenes used by the parity tests: small enough for the oracle to finish in
seconds, covering every option the HIP path implements."""
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


def spheres_wide_gaussian():
    # A 5x5 filter footprint (DiscreteShift 2) with 3x3 samples and a frame
    # that is not a whole number of 8x8 pixel tiles.
    return scenes.spheres_scene(seed=19, n=40, res=(150, 98), samples=(3, 3), filt=("gaussian", 4.0, 4.0))


CASES = {
    "spheres_config": spheres_config,
    "patches_small": patches_small,
    "patches_8x8": patches_8x8,
    "spheres_box_2x2": spheres_box_2x2,
    "patches_crop_odd": patches_crop_odd,
    "patches_gridsampler": patches_gridsampler,
    "spheres_exposure_mitchell": spheres_exposure_mitchell,
    "spheres_wide_gaussian": spheres_wide_gaussian,
}
