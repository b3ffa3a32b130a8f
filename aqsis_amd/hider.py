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


class RhTuning(C.Structure):
    _fields_ = [("mode", C.c_int32), ("lds_records", C.c_int32), ("blocks_per_wave", C.c_int32),
                ("split_cost", C.c_uint32), ("frame_ms", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


MODES = {"grid": 0, "plain": 1, "batched": 2}

# display request modes (include/aqsis_hider.h RH_MODE_*)
MODE_RGB, MODE_A, MODE_Z = 1, 2, 4


def mode_bits(mode):
    """RH_MODE_* bits of a RiDisplay mode string: "rgb", then "a", then "z",
    each optional and matched as a prefix in that order (ri.cpp:947-964)."""
    bits, i = 0, 0
    if mode.startswith("rgb"):
        bits, i = MODE_RGB, 3
    if mode.startswith("a", i):
        bits, i = bits | MODE_A, i + 1
    if mode.startswith("z", i):
        bits |= MODE_Z
    return bits

EXPORTS = ["rh_options_default", "rh_create", "rh_post_surfaces", "rh_clear_surfaces", "rh_render",
           "rh_render_device", "rh_render_device_noalloc", "rh_render_stages", "rh_frame_size", "rh_read_channels", "rh_read_rgba8", "rh_display", "rh_tiff_write",
           "rh_display_format", "rh_read_display", "rh_format_display", "rh_display_as", "rh_tiff_write_typed",
           "rh_select_format", "rh_read_display_request",
           "rh_read_samples",
           "rh_get_stats", "rh_destroy", "rh_last_error", "rh_estimate", "rh_sample_positions",
           "rh_set_tuning", "rh_get_tuning", "rh_autotune"]


# Display value types (RH_DSPY_*, the PkDspy* codes) and their numpy dtypes.
DSPY = {"float32": 1, "uint32": 2, "int32": 3, "uint16": 4, "int16": 5, "uint8": 6, "int8": 7}
DSPY_DTYPE = {1: np.float32, 2: np.uint32, 3: np.int32, 4: np.uint16, 5: np.int16, 6: np.uint8, 7: np.int8}
# The file driver receives signed types as the unsigned type of their width.
FILE_TYPE = {1: 1, 2: 2, 3: 2, 4: 4, 5: 4, 6: 6, 7: 6}

DSPY_DATA_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p)


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
        L.rh_render_device_noalloc.argtypes = [C.c_void_p]
        L.rh_render_stages.argtypes = [C.c_void_p]
        L.rh_frame_size.argtypes = [C.c_void_p] + [P(C.c_int32)] * 5
        L.rh_read_channels.argtypes = [C.c_void_p, C.c_void_p]
        L.rh_read_rgba8.argtypes = [C.c_void_p, C.c_void_p]
        L.rh_read_samples.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.rh_display.argtypes = [C.c_void_p, DSPY_DATA_FN, C.c_void_p, C.c_int32]
        L.rh_tiff_write.argtypes = [C.c_char_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                    C.c_void_p, C.c_void_p, C.c_char_p]
        L.rh_display_format.argtypes = [P(RhOptions), P(C.c_int32), P(C.c_int32)]
        L.rh_read_display.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.rh_format_display.argtypes = [P(RhOptions), C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p]
        L.rh_display_as.argtypes = [C.c_void_p, DSPY_DATA_FN, C.c_void_p, C.c_int32, C.c_int32]
        L.rh_select_format.argtypes = [C.c_float, C.c_float, C.c_float, P(C.c_int32)]
        L.rh_read_display_request.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]
        L.rh_tiff_write_typed.argtypes = [C.c_char_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                          C.c_int32, C.c_void_p, C.c_void_p, C.c_char_p]
        L.rh_get_stats.argtypes = [C.c_void_p, P(RhStats)]
        L.rh_destroy.argtypes = [C.c_void_p]
        L.rh_destroy.restype = None
        L.rh_last_error.argtypes = []
        L.rh_last_error.restype = C.c_char_p
        L.rh_estimate.argtypes = [P(RhOptions), P(RhPrim), C.c_int32, P(RhFootprint)]
        L.rh_sample_positions.argtypes = [P(RhOptions), C.c_void_p, C.c_void_p]
        L.rh_set_tuning.argtypes = [C.c_void_p, P(RhTuning)]
        L.rh_get_tuning.argtypes = [C.c_void_p, P(RhTuning)]
        L.rh_autotune.argtypes = [C.c_void_p, P(RhTuning)]
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
        # the crop window's first pixel, ceil(res * crop) in single precision
        f32 = np.float32
        self.origin = (min(max(int(np.ceil(f32(opts.xres) * f32(opts.crop[0]))), 0), opts.xres),
                       min(max(int(np.ceil(f32(opts.yres) * f32(opts.crop[2]))), 0), opts.yres))
        self.full_size = (opts.xres, opts.yres)
        self.display_type = display_format(opts)[0]

    def post_surfaces(self, prims):
        arr, n = _prims(prims)
        _check(lib().rh_post_surfaces(self._ctx, arr, n))

    def clear_surfaces(self):
        _check(lib().rh_clear_surfaces(self._ctx))

    def render(self):
        _check(lib().rh_render(self._ctx))

    def render_device(self):
        _check(lib().rh_render_device(self._ctx))

    def render_device_noalloc(self):
        """rh_render_device with no device allocation (scratch array of the GPU)."""
        _check(lib().rh_render_device_noalloc(self._ctx))

    def render_stages(self):
        """One frame issued kernel by kernel with per-stage device times."""
        _check(lib().rh_render_stages(self._ctx))

    def channels(self):
        out = np.zeros((self.height, self.width, 9), np.float32)
        _check(lib().rh_read_channels(self._ctx, _ptr(out)))
        return out

    def rgba8(self):
        out = np.zeros((self.height, self.width, 4), np.uint8)
        _check(lib().rh_read_rgba8(self._ctx, _ptr(out)))
        return out

    def display_values(self, type=None):
        """The frame's display values (height, width, 4: r, g, b, a) as
        FormatBucketForDisplay gives them, of RH_DSPY_* `type` (default the
        type the Quantize settings select) (rh_read_display)."""
        t = self.display_type if type is None else (DSPY[type] if isinstance(type, str) else int(type))
        out = np.zeros((self.height, self.width, 4), DSPY_DTYPE[t])
        _check(lib().rh_read_display(self._ctx, t, _ptr(out)))
        return out

    def display_request_values(self, mode, quantize=(0.0, 0.0, 0.0, 0.0), index=0, type=None):
        """A display request's values (height, width, k) of mode `mode`
        (RH_MODE_* bits or a mode string such as "z" or "rgbaz"), channels in
        the request's format order a, r, g, b, z, quantised with `quantize`
        (one, min, max, dither: the request's settings, Quantize "z" for a z
        mode; the default is Quantize "z"'s default, floats) and dither draws
        of display request `index` (rh_read_display_request); `type`
        defaults to the one selectDataFormat picks for the settings."""
        m = mode_bits(mode) if isinstance(mode, str) else int(mode)
        q = np.ascontiguousarray(quantize, np.float32)
        if type is None:
            t = C.c_int32()
            _check(lib().rh_select_format(float(q[0]), float(q[1]), float(q[2]), C.byref(t)))
            t = t.value
        else:
            t = DSPY[type] if isinstance(type, str) else int(type)
        k = (1 if m & MODE_A else 0) + (3 if m & MODE_RGB else 0) + (1 if m & MODE_Z else 0)
        out = np.zeros((self.height, self.width, max(k, 1)), DSPY_DTYPE.get(t, np.uint8))
        _check(lib().rh_read_display_request(self._ctx, m, _ptr(q), index, t, _ptr(out)))
        return out

    def display(self, data, scanline_order=False, type=None):
        """Deliver the display values to `data(xmin, xmaxplus1, ymin,
        ymaxplus1, entry_size, bytes)`, raster coordinates, as a display
        driver's DspyImageData receives them (rh_display / rh_display_as); a
        true return value from `data` stops the delivery."""
        def cb(_image, x0, x1, y0, y1, es, ptr):
            n = (x1 - x0) * (y1 - y0) * es
            return 1 if data(x0, x1, y0, y1, es, C.string_at(ptr, n)) else 0
        fn = DSPY_DATA_FN(cb)
        order = 1 if scanline_order else 0
        if type is None:
            _check(lib().rh_display(self._ctx, fn, None, order))
        else:
            t = DSPY[type] if isinstance(type, str) else int(type)
            _check(lib().rh_display_as(self._ctx, fn, None, order, t))

    def write_tiff(self, path, compression="lzw", scanline_order=False, channels=4):
        """Write the frame to a TIFF file as the "file" display driver does:
        the driver asks for the unsigned type of the frame's display type,
        fills its image buffer from the delivery (DspyImageData, raster
        coordinates less the crop window's origin), and writes it with
        rh_tiff_write_typed; channels 4 (rgba) or 3 (rgb).  A cropped frame
        carries its origin and the full image size."""
        t = FILE_TYPE[self.display_type]
        dt = np.dtype(DSPY_DTYPE[t])
        img = np.zeros((self.height, self.width, 4), dt)
        ox, oy = self.origin

        def data(x0, x1, y0, y1, es, buf):
            img[y0 - oy:y1 - oy, x0 - ox:x1 - ox] = np.frombuffer(buf, dt).reshape(y1 - y0, x1 - x0, 4)
            return False
        self.display(data, scanline_order, type=t)
        cropped = (self.width, self.height) != self.full_size
        write_tiff(path, img[..., :channels], compression, origin=self.origin if cropped else None,
                   full_size=self.full_size if cropped else None)

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

    def tuning(self):
        t = RhTuning()
        _check(lib().rh_get_tuning(self._ctx, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, mode=None, lds_records=None, blocks_per_wave=None, split_cost=None):
        """Device layout of the following frames (see rh_tuning); unspecified
        fields keep their current values.  mode: "grid", "plain", "batched"."""
        t = RhTuning()
        _check(lib().rh_get_tuning(self._ctx, C.byref(t)))
        if mode is not None:
            t.mode = MODES[mode] if isinstance(mode, str) else int(mode)
        if lds_records is not None:
            t.lds_records = lds_records
        if blocks_per_wave is not None:
            t.blocks_per_wave = blocks_per_wave
        if split_cost is not None:
            t.split_cost = split_cost
        _check(lib().rh_set_tuning(self._ctx, C.byref(t)))

    def autotune(self):
        """Pick the fastest layout (static grid or batched setting) for the
        posted scene (renders frames)."""
        t = RhTuning()
        _check(lib().rh_autotune(self._ctx, C.byref(t)))
        return t.as_dict()

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


TIFF_COMPRESSION = {"none": 1, "lzw": 5}


def write_tiff(path, pixels, compression="lzw", origin=None, full_size=None, description=None):
    """Write a (height, width, 3 or 4) image to `path` as the "file" display
    driver does: uint8 through its integer branch (rh_tiff_write); uint16,
    uint32 and float32 through its other branch (rh_tiff_write_typed)."""
    a = np.ascontiguousarray(pixels)
    types = {np.dtype(np.uint8): 6, np.dtype(np.uint16): 4, np.dtype(np.uint32): 2, np.dtype(np.float32): 1}
    if a.ndim != 3 or a.dtype not in types:
        raise HiderError("write_tiff: pixels must be (height, width, channels) of uint8, uint16, uint32 or float32")
    o = None if origin is None else (C.c_int32 * 2)(*origin)
    f = None if full_size is None else (C.c_int32 * 2)(*full_size)
    d = None if description is None else description.encode()
    if a.dtype == np.uint8:
        _check(lib().rh_tiff_write(os.fsencode(path), _ptr(a), a.shape[1], a.shape[0], a.shape[2],
                                   TIFF_COMPRESSION[compression], o, f, d))
    else:
        _check(lib().rh_tiff_write_typed(os.fsencode(path), _ptr(a), a.shape[1], a.shape[0], a.shape[2],
                                         types[a.dtype], TIFF_COMPRESSION[compression], o, f, d))


def format_display(opts, channels, dither, type=None):
    """Host-only FormatBucketForDisplay of (h, w, 9) channels with (h, w)
    dither draws: (h, w, 4) values of RH_DSPY_* `type` (default the type the
    Quantize settings select) (rh_format_display)."""
    t = display_format(opts)[0] if type is None else (DSPY[type] if isinstance(type, str) else int(type))
    ch = np.ascontiguousarray(channels, np.float32)
    d = np.ascontiguousarray(dither, np.float32)
    if ch.ndim != 3 or ch.shape[2] != 9 or d.shape != ch.shape[:2]:
        raise HiderError("format_display: channels (h, w, 9) and dither (h, w) expected")
    out = np.zeros(ch.shape[:2] + (4,), DSPY_DTYPE[t])
    _check(lib().rh_format_display(C.byref(opts), _ptr(ch), _ptr(d), ch.shape[0] * ch.shape[1], t, _ptr(out)))
    return out


def display_format(opts):
    """(RH_DSPY_* type, bytes per value) the Quantize settings select."""
    t, b = C.c_int32(), C.c_int32()
    _check(lib().rh_display_format(C.byref(opts), C.byref(t), C.byref(b)))
    return t.value, b.value


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


class RhServerOptions(C.Structure):
    _fields_ = [("slots", C.c_int32), ("schedule_threads", C.c_int32), ("memory_budget", C.c_uint64)]


class RhJobInfo(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int32), ("state", C.c_int32), ("width", C.c_int32),
                ("height", C.c_int32), ("device_bytes", C.c_uint64), ("spl_count", C.c_uint64),
                ("ms_queue", C.c_double), ("ms_run", C.c_double), ("ms_total", C.c_double),
                ("ms_device", C.c_double), ("ms_setup", C.c_double), ("ms_upload", C.c_double),
                ("ms_render", C.c_double), ("ms_readback", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


SERVER_EXPORTS = ["rh_server_create", "rh_server_submit", "rh_server_wait", "rh_server_poll", "rh_server_read",
                  "rh_server_release", "rh_server_memory", "rh_server_destroy", "rh_server_last_error"]


def _server_lib():
    L = lib()
    if not getattr(L, "_server_bound", False):
        P = C.POINTER
        L.rh_server_create.argtypes = [P(RhServerOptions), P(C.c_void_p)]
        L.rh_server_submit.argtypes = [C.c_void_p, P(RhOptions), P(RhPrim), C.c_int32, P(C.c_uint64)]
        L.rh_server_wait.argtypes = [C.c_void_p, C.c_uint64, P(RhJobInfo)]
        L.rh_server_poll.argtypes = [C.c_void_p, P(RhJobInfo), C.c_int32, P(C.c_int32)]
        L.rh_server_read.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]
        L.rh_server_release.argtypes = [C.c_void_p, C.c_uint64]
        L.rh_server_memory.argtypes = [C.c_void_p, P(C.c_uint64), P(C.c_uint64)]
        L.rh_server_destroy.argtypes = [C.c_void_p]
        L.rh_server_destroy.restype = None
        L.rh_server_last_error.argtypes = []
        L.rh_server_last_error.restype = C.c_char_p
        for name in SERVER_EXPORTS[:-2]:
            getattr(L, name).restype = C.c_int32
        L._server_bound = True
    return L


def _scheck(status):
    if status != 0:
        raise HiderError(f"status {status}: {_server_lib().rh_server_last_error().decode()}")


class Server:
    """Serving loop (rh_server_*): jobs of any frame shape run concurrently on
    one GPU, admitted by device-memory footprint, contexts resident by shape."""

    def __init__(self, slots=4, memory_budget=64 << 30, schedule_threads=0):
        self._srv = C.c_void_p()
        o = RhServerOptions(slots, schedule_threads, memory_budget)
        _scheck(_server_lib().rh_server_create(C.byref(o), C.byref(self._srv)))

    def submit(self, opts, prims):
        arr, n = _prims(prims)
        jid = C.c_uint64()
        _scheck(_server_lib().rh_server_submit(self._srv, C.byref(opts), arr, n, C.byref(jid)))
        return jid.value

    def wait(self, jid):
        info = RhJobInfo()
        _scheck(_server_lib().rh_server_wait(self._srv, jid, C.byref(info)))
        return info.as_dict()

    def poll(self, max_jobs=256):
        infos = (RhJobInfo * max_jobs)()
        n = C.c_int32()
        _scheck(_server_lib().rh_server_poll(self._srv, infos, max_jobs, C.byref(n)))
        return [infos[i].as_dict() for i in range(n.value)]

    def result(self, jid, width, height):
        ch = np.zeros((height, width, 9), np.float32)
        rgba = np.zeros((height, width, 4), np.uint8)
        _scheck(_server_lib().rh_server_read(self._srv, jid, _ptr(ch), _ptr(rgba)))
        return ch, rgba

    def release(self, jid):
        _scheck(_server_lib().rh_server_release(self._srv, jid))

    def memory(self):
        r, hw = C.c_uint64(), C.c_uint64()
        _scheck(_server_lib().rh_server_memory(self._srv, C.byref(r), C.byref(hw)))
        return r.value, hw.value

    def close(self):
        if self._srv:
            _server_lib().rh_server_destroy(self._srv)
            self._srv = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
