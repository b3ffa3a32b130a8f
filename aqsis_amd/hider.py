"""Python binding of the HIP hider (libaqsis_hider.so, include/aqsis_hider.h).

`Hider` mirrors the reference's frame driver: construction is
CqImageBuffer::SetImage with the cached options, `post_surfaces` is
CqImageBuffer::PostSurface, `render` is CqImageBuffer::RenderImage and the
read-back methods return what the display manager would receive.  There is
no CPU fallback: if the library or a GPU is missing, construction raises.
"""
import ctypes as C
import os

import numpy as np

from .abi import RhOptions, RhPrim, RhStats

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libaqsis_hider.so")
_lib = None


class HiderError(RuntimeError):
    pass


class RhFootprint(C.Structure):
    _fields_ = [("surfaces", C.c_uint64), ("grids", C.c_uint64), ("micropolygons", C.c_uint64),
                ("vertices", C.c_uint64), ("list_capacity", C.c_uint64), ("device_bytes", C.c_uint64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


EXPORTS = ["rh_options_default", "rh_create", "rh_post_surfaces", "rh_clear_surfaces", "rh_render",
           "rh_render_device", "rh_frame_size", "rh_read_channels", "rh_read_rgba8", "rh_read_samples",
           "rh_get_stats", "rh_destroy", "rh_last_error", "rh_estimate", "rh_sample_positions"]


def lib():
    """Load the in-tree library (built by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HiderError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.rh_options_default.argtypes = [P(RhOptions)]
        L.rh_options_default.restype = None
        L.rh_create.argtypes = [P(RhOptions), P(C.c_void_p)]
        L.rh_post_surfaces.argtypes = [C.c_void_p, P(RhPrim), C.c_int32]
        L.rh_clear_surfaces.argtypes = [C.c_void_p]
        L.rh_render.argtypes = [C.c_void_p]
        L.rh_render_device.argtypes = [C.c_void_p]
        L.rh_frame_size.argtypes = [C.c_void_p] + [P(C.c_int32)] * 5
        L.rh_read_channels.argtypes = [C.c_void_p, C.c_void_p]
        L.rh_read_rgba8.argtypes = [C.c_void_p, C.c_void_p]
        L.rh_read_samples.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.rh_get_stats.argtypes = [C.c_void_p, P(RhStats)]
        L.rh_destroy.argtypes = [C.c_void_p]
        L.rh_destroy.restype = None
        L.rh_last_error.argtypes = []
        L.rh_last_error.restype = C.c_char_p
        L.rh_estimate.argtypes = [P(RhOptions), P(RhPrim), C.c_int32, P(RhFootprint)]
        L.rh_sample_positions.argtypes = [P(RhOptions), C.c_void_p, C.c_void_p]
        for name in EXPORTS:
            if name not in ("rh_options_default", "rh_destroy", "rh_last_error"):
                getattr(L, name).restype = C.c_int32
        _lib = L
    return _lib


def _check(status):
    if status != 0:
        raise HiderError(f"rh status {status}: {lib().rh_last_error().decode()}")


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _prims(prims):
    from .scenes import prim_array
    if isinstance(prims, C.Array):
        return prims, len(prims)
    arr = prim_array(list(prims))
    return arr, len(arr)


class Hider:
    """One frame shape on the GPU; surfaces can be posted, rendered and read."""

    def __init__(self, opts):
        self._ctx = C.c_void_p()
        _check(lib().rh_create(C.byref(opts), C.byref(self._ctx)))
        v = [C.c_int32() for _ in range(5)]
        _check(lib().rh_frame_size(self._ctx, *[C.byref(x) for x in v]))
        self.width, self.height, self.ext_width, self.ext_height, self.spp = (x.value for x in v)

    def post_surfaces(self, prims):
        arr, n = _prims(prims)
        _check(lib().rh_post_surfaces(self._ctx, arr, n))

    def clear_surfaces(self):
        _check(lib().rh_clear_surfaces(self._ctx))

    def render(self):
        _check(lib().rh_render(self._ctx))

    def render_device(self):
        _check(lib().rh_render_device(self._ctx))

    def channels(self):
        out = np.zeros((self.height, self.width, 9), np.float32)
        _check(lib().rh_read_channels(self._ctx, _ptr(out)))
        return out

    def rgba8(self):
        out = np.zeros((self.height, self.width, 4), np.uint8)
        _check(lib().rh_read_rgba8(self._ctx, _ptr(out)))
        return out

    def samples(self):
        n = self.ext_width * self.ext_height * self.spp
        ids = np.zeros(n, np.uint32)
        depth = np.zeros(n, np.float32)
        rgbaz = np.zeros((n, 7), np.float32)
        _check(lib().rh_read_samples(self._ctx, _ptr(ids), _ptr(depth), _ptr(rgbaz)))
        return ids, depth, rgbaz

    def stats(self):
        st = RhStats()
        _check(lib().rh_get_stats(self._ctx, C.byref(st)))
        return st.as_dict()

    def close(self):
        if self._ctx:
            lib().rh_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def estimate(opts, prims):
    """Host-only schedule summary and device footprint of one frame."""
    arr, n = _prims(prims)
    fp = RhFootprint()
    _check(lib().rh_estimate(C.byref(opts), arr, n, C.byref(fp)))
    return fp.as_dict()


def sample_positions(opts, ext_shape, display_shape):
    """Host-only: jittered sample positions (ext pixels x spp x 2) and the
    dither draw of every display pixel."""
    ew, eh, spp = ext_shape
    w, h = display_shape
    xy = np.zeros((eh, ew, spp, 2), np.float32)
    d = np.zeros((h, w), np.float32)
    _check(lib().rh_sample_positions(C.byref(opts), _ptr(xy), _ptr(d)))
    return xy, d
