"""ctypes mirror of include/aqsis_hider.h.

The structures here are byte-for-byte the C structs of the hider ABI; the
same layouts are read by the HIP library (libaqsis_hider.so) and by any
other implementation of the ABI.
"""
import ctypes as C

RH_OK = 0

RH_FILTER_BOX = 0
RH_FILTER_TRIANGLE = 1
RH_FILTER_CATMULLROM = 2
RH_FILTER_GAUSSIAN = 3
RH_FILTER_SINC = 4
RH_FILTER_DISK = 5
RH_FILTER_BESSEL = 6
RH_FILTER_MITCHELL = 7

FILTERS = {
    "box": RH_FILTER_BOX, "triangle": RH_FILTER_TRIANGLE,
    "catmull-rom": RH_FILTER_CATMULLROM, "gaussian": RH_FILTER_GAUSSIAN,
    "sinc": RH_FILTER_SINC, "disk": RH_FILTER_DISK, "bessel": RH_FILTER_BESSEL,
    "mitchell": RH_FILTER_MITCHELL,
}

RH_PRIM_SPHERE = 1
RH_PRIM_BICUBIC = 2
RH_PRIM_BILINEAR = 3

# rh_options.depth_filter: Hider "hidden" "depthfilter" (optioncache.cpp:96-110)
RH_DEPTH_MIN = 0
RH_DEPTH_MIDPOINT = 1
RH_DEPTH_MAX = 2
RH_DEPTH_AVERAGE = 3
DEPTH_FILTERS = {"min": RH_DEPTH_MIN, "midpoint": RH_DEPTH_MIDPOINT, "max": RH_DEPTH_MAX,
                 "average": RH_DEPTH_AVERAGE}

# rh_prim.matte: RiMatte's value as CacheGridInfo reads it (micropolygon.cpp:49-55)
RH_MATTE_NONE = 0
RH_MATTE = 1
RH_MATTE_ALPHA = 2

FLT_EPSILON = 1.1920928955078125e-07
FLT_MAX = 3.4028234663852886e38


class RhOptions(C.Structure):
    _fields_ = [
        ("xres", C.c_int32), ("yres", C.c_int32),
        ("pixel_aspect", C.c_float),
        ("crop", C.c_float * 4),
        ("fov", C.c_float),
        ("clip_near", C.c_float), ("clip_far", C.c_float),
        ("xsamples", C.c_int32), ("ysamples", C.c_int32),
        ("filter", C.c_int32),
        ("filter_xwidth", C.c_float), ("filter_ywidth", C.c_float),
        ("exposure_gain", C.c_float), ("exposure_gamma", C.c_float),
        ("quantize_one", C.c_float), ("quantize_min", C.c_float),
        ("quantize_max", C.c_float), ("quantize_dither", C.c_float),
        ("shutter_open", C.c_float), ("shutter_close", C.c_float),
        ("bucket_x", C.c_int32), ("bucket_y", C.c_int32),
        ("eyesplits", C.c_int32),
        ("jitter", C.c_int32),
        ("num_displays", C.c_int32),
        ("zthreshold", C.c_float * 3),
        ("dof", C.c_float * 3),
        ("depth_filter", C.c_int32),
        ("display_z", C.c_int32),
    ]


class RhPrim(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("nkeys", C.c_int32),
        ("shading_rate", C.c_float),
        ("Cs", C.c_float * 3),
        ("Os", C.c_float * 3),
        ("key_time", C.c_float * 2),
        ("sphere", C.c_float * 4),
        ("xform", (C.c_float * 16) * 2),
        ("P", (C.c_float * 3) * 16),
        ("P1", (C.c_float * 3) * 16),
        ("no_cull_hidden", C.c_int32),
        ("focus_factor", C.c_float),
        ("shading_interpolation", C.c_int32),
        ("matte", C.c_int32),
    ]


class RhStats(C.Structure):
    _fields_ = [
        ("spl_count", C.c_uint64), ("spl_bound_hits", C.c_uint64),
        ("spl_hits", C.c_uint64), ("inside", C.c_uint64),
        ("inside_nonlinear", C.c_uint64), ("grids", C.c_uint64),
        ("micropolygons", C.c_uint64), ("mp_bucket_refs", C.c_uint64),
        ("surfaces", C.c_uint64), ("occlusion_culled", C.c_uint64),
        ("ms_schedule", C.c_double), ("ms_upload", C.c_double),
        ("ms_dice", C.c_double), ("ms_bin", C.c_double), ("ms_sample", C.c_double),
        ("ms_filter", C.c_double), ("ms_device", C.c_double), ("ms_composite", C.c_double),
        ("cull_probes", C.c_uint64), ("cull_passes", C.c_uint64), ("ms_cull", C.c_double),
        ("ms_forest", C.c_double), ("ms_simulate", C.c_double), ("ms_cull_depth", C.c_double),
        ("ms_cull_probes", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def default_options(**kw):
    """The reference's option defaults (libs/core/options.cpp:273-310,
    optioncache.cpp:48-117), overridden by keyword arguments."""
    o = RhOptions()
    o.xres, o.yres = 640, 480
    o.pixel_aspect = 1.0
    o.crop[:] = [0.0, 1.0, 0.0, 1.0]
    o.fov = 90.0
    o.clip_near, o.clip_far = FLT_EPSILON, FLT_MAX
    o.xsamples, o.ysamples = 2, 2
    o.filter = RH_FILTER_GAUSSIAN
    o.filter_xwidth, o.filter_ywidth = 2.0, 2.0
    o.exposure_gain, o.exposure_gamma = 1.0, 1.0
    o.quantize_one, o.quantize_min, o.quantize_max, o.quantize_dither = 255.0, 0.0, 255.0, 0.5
    o.shutter_open, o.shutter_close = 0.0, 0.0
    o.bucket_x, o.bucket_y = 16, 16
    o.eyesplits = 10
    o.jitter = 1
    o.num_displays = 1
    o.zthreshold[:] = [1.0, 1.0, 1.0]
    o.dof[:] = [FLT_MAX, FLT_MAX, FLT_MAX]
    for k, v in kw.items():
        if k == "filter" and isinstance(v, str):
            v = FILTERS[v]
        if k in ("crop", "zthreshold", "dof"):
            getattr(o, k)[:] = list(v)
        else:
            setattr(o, k, v)
    return o
