"""ORACLE loader -- test infrastructure only.

Loads oracle/_build/liboracle.so (built by `make -C oracle`) and wraps its C
API.  Only tests/, the smoke check and bench.py's CPU-baseline leg use this
module; the product never does.
"""
import ctypes as C
import os

import numpy as np

from aqsis_amd.abi import RhOptions, RhPrim, RhStats

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.oracle_frame_size.argtypes = [C.POINTER(RhOptions)] + [C.POINTER(C.c_int32)] * 5
        L.oracle_render.argtypes = [C.POINTER(RhOptions), C.POINTER(RhPrim), C.c_int32, C.c_int32,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.POINTER(RhStats)]
        L.oracle_mt_outputs.argtypes = [C.c_uint32, C.c_int32, C.c_void_p]
        L.oracle_random_floats.argtypes = [C.c_uint32, C.c_int32, C.c_void_p]
        L.oracle_round_f.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_inv_bilinear.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
        L.oracle_bilerp2.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
        L.oracle_tree_index.argtypes = [C.c_int32] * 4
        L.oracle_tree_index.restype = C.c_int32
        L.oracle_can_cull.argtypes = [C.c_int32] * 6 + [C.c_void_p, C.c_void_p]
        L.oracle_can_cull.restype = C.c_int32
        L.oracle_filter.argtypes = [C.c_int32, C.c_float, C.c_float, C.c_float, C.c_float]
        L.oracle_filter.restype = C.c_float
        L.oracle_jitter_tables.argtypes = [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_mp_sample.argtypes = [C.c_void_p, C.c_float, C.c_float, C.POINTER(C.c_float)]
        L.oracle_mp_sample.restype = C.c_int32
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def frame_size(opts):
    v = [C.c_int32() for _ in range(5)]
    lib().oracle_frame_size(C.byref(opts), *[C.byref(x) for x in v])
    return tuple(x.value for x in v)


def render(opts, prims, occlusion_cull=False, samples=True):
    """Render on the CPU; returns a dict of numpy arrays and the stats."""
    from aqsis_amd.scenes import prim_array
    w, h, ew, eh, spp = frame_size(opts)
    arr = prim_array(prims)
    ch = np.zeros((h, w, 9), np.float32)
    rgba = np.zeros((h, w, 4), np.uint8)
    ns = ew * eh * spp
    ids = np.zeros(ns, np.uint32) if samples else None
    depth = np.zeros(ns, np.float32) if samples else None
    rgbaz = np.zeros((ns, 7), np.float32) if samples else None
    st = RhStats()
    rc = lib().oracle_render(C.byref(opts), arr, len(prims), 1 if occlusion_cull else 0,
                             _ptr(ch), _ptr(rgba), _ptr(ids), _ptr(depth), _ptr(rgbaz), C.byref(st))
    if rc != 0:
        raise ValueError(f"oracle_render: unsupported scene (code {rc}: depth of field with motion keys)")
    return {"channels": ch, "rgba8": rgba, "mp_id": ids, "depth": depth, "rgbaz": rgbaz,
            "stats": st.as_dict(), "ext": (ew, eh, spp)}


def sample_positions(opts, display=0):
    """The sample positions (eh, ew, spp, 2) and the (h, w) dither draws of
    display request `display`."""
    w, h, ew, eh, spp = frame_size(opts)
    L = lib()
    L.oracle_sample_positions_of.argtypes = [C.POINTER(RhOptions), C.c_int32, C.c_void_p, C.c_void_p]
    xy = np.zeros((eh, ew, spp, 2), np.float32)
    d = np.zeros((h, w), np.float32)
    L.oracle_sample_positions_of(C.byref(opts), display, _ptr(xy), _ptr(d))
    return xy, d


def sample_times(opts):
    w, h, ew, eh, spp = frame_size(opts)
    L = lib()
    L.oracle_sample_times.argtypes = [C.POINTER(RhOptions), C.c_void_p]
    t = np.zeros((eh, ew, spp), np.float32)
    L.oracle_sample_times(C.byref(opts), _ptr(t))
    return t


def sample_lens(opts):
    """Lens offset of every sample of the extended image (eh, ew, spp, 2) and
    the lens cell each sample is visited through (eh, ew, spp)."""
    w, h, ew, eh, spp = frame_size(opts)
    L = lib()
    L.oracle_sample_lens.argtypes = [C.POINTER(RhOptions), C.c_void_p, C.c_void_p]
    xy = np.zeros((eh, ew, spp, 2), np.float32)
    cell = np.zeros((eh, ew, spp), np.int32)
    L.oracle_sample_lens(C.byref(opts), _ptr(xy), _ptr(cell))
    return xy, cell


def dof(opts, z):
    """The frame's circle of confusion at camera depths z (n, 2), its lens-cell
    bounds (spp, 4) and MinCoCForBound of consecutive depth pairs (n - 1)."""
    w, h, ew, eh, spp = frame_size(opts)
    L = lib()
    L.oracle_dof.argtypes = [C.POINTER(RhOptions), C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    z = np.ascontiguousarray(z, np.float32)
    coc = np.zeros((len(z), 2), np.float32)
    cells = np.zeros((spp, 4), np.float32)
    mn = np.zeros(max(len(z) - 1, 1), np.float32)
    if L.oracle_dof(C.byref(opts), len(z), _ptr(z), _ptr(coc), _ptr(cells), _ptr(mn)) != 0:
        raise ValueError("the frame has no depth of field")
    return coc, cells, mn[:len(z) - 1]


DISPLAY_DTYPE = {1: np.float32, 2: np.uint32, 3: np.int32, 4: np.uint16, 5: np.int16, 6: np.uint8, 7: np.int8}


def display_type(opts):
    """selectDataFormat's PkDspy* code for the frame's rgba display."""
    L = lib()
    L.oracle_display_type.argtypes = [C.POINTER(RhOptions)]
    L.oracle_display_type.restype = C.c_int32
    return L.oracle_display_type(C.byref(opts))


def display_values(opts, channels, dither, type=None):
    """FormatBucketForDisplay's (h, w, 4) values of PkDspy* `type` (default
    the frame's) from the (h, w, 9) channels and (h, w) dither draws."""
    L = lib()
    L.oracle_display_values.argtypes = [C.POINTER(RhOptions), C.c_void_p, C.c_void_p, C.c_int64, C.c_int32,
                                        C.c_void_p]
    t = display_type(opts) if type is None else type
    ch = np.ascontiguousarray(channels, np.float32)
    d = np.ascontiguousarray(dither, np.float32)
    h, w = ch.shape[:2]
    assert d.shape == (h, w)
    out = np.zeros((h, w, 4), DISPLAY_DTYPE[t])
    L.oracle_display_values(C.byref(opts), _ptr(ch), _ptr(d), h * w, t, _ptr(out))
    return out


MODE_CHANNELS = {1: 3, 2: 1, 4: 1}


def display_request_values(quantize, mode, channels, dither, type):
    """A display request's (h, w, k) values of PkDspy* `type`: mode bits (1
    rgb, 2 a, 4 z) in the request's format order a, r, g, b, z, quantised
    with (one, min, max, dither)."""
    L = lib()
    L.oracle_display_request_values.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int64,
                                                C.c_int32, C.c_void_p]
    q = np.ascontiguousarray(quantize, np.float32)
    ch = np.ascontiguousarray(channels, np.float32)
    d = np.ascontiguousarray(dither, np.float32)
    h, w = ch.shape[:2]
    k = sum(n for bit, n in MODE_CHANNELS.items() if mode & bit)
    out = np.zeros((h, w, k), DISPLAY_DTYPE[type])
    L.oracle_display_request_values(_ptr(q), mode, _ptr(ch), _ptr(d), h * w, type, _ptr(out))
    return out


def dice(prim, u, v):
    L = lib()
    L.oracle_dice.argtypes = [C.POINTER(RhPrim), C.c_int32, C.c_int32, C.c_void_p]
    out = np.zeros(((v + 1) * (u + 1), 3), np.float32)
    L.oracle_dice(C.byref(prim), u, v, _ptr(out))
    return out
