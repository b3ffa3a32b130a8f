"""A RIB front end for the hider: the subset of the RenderMan Interface
Bytestream that drives the hot path, read into rh_options and camera-space
rh_prim records.

Requests (libs/core/api/ri.cpp; parsing follows libs/riutil/ribparser.cpp's
token rules: strings in double quotes, numbers, [ ] arrays, # comments):

  Format, PixelSamples, PixelFilter, ShadingRate, Exposure, Quantize "rgba",
  Projection "perspective" "fov", Clipping, CropWindow, Shutter,
  Option "limits" "bucketsize" / "eyesplits" / "zthreshold",
  Hider "hidden" "jitter", Display (ignored), FrameBegin/End, WorldBegin/End,
  AttributeBegin/End, TransformBegin/End, Identity, Transform,
  ConcatTransform, Translate, Rotate, Scale, Color, Opacity,
  Surface "constant", Sphere, Patch "bicubic" / "bilinear" "P".

Transformations use RenderMan's row-vector convention: a request's matrix
M is applied before the current one (CTM <- M * CTM); transformations before
WorldBegin make the camera transformation, and surfaces are posted in camera
space.  Anything else raises RibError naming the request and line.
"""
import math
import re

import numpy as np

from .abi import FILTERS, default_options
from . import scenes


class RibError(ValueError):
    pass


_TOKEN = re.compile(r'"([^"]*)"|(\[)|(\])|([-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?)|([A-Za-z_][A-Za-z0-9_]*)|(#[^\n]*)|(\S)')


def _tokens(text):
    line = 1
    pos = 0
    for m in _TOKEN.finditer(text):
        line += text.count("\n", pos, m.start())
        pos = m.start()
        s, lb, rb, num, word, comment, bad = m.groups()
        if comment is not None:
            continue
        if bad is not None:
            raise RibError(f"line {line}: unexpected character {bad!r}")
        if s is not None:
            yield line, "str", s
        elif lb:
            yield line, "[", None
        elif rb:
            yield line, "]", None
        elif num is not None:
            yield line, "num", float(num)
        else:
            yield line, "word", word


def _requests(text):
    """(line, name, args): args are floats, strings or lists of either."""
    name, args, line0 = None, [], 0
    arr = None
    for line, kind, val in _tokens(text):
        if kind == "word":
            if arr is not None:
                raise RibError(f"line {line}: request {val} inside an array")
            if name is not None:
                yield line0, name, args
            name, args, line0 = val, [], line
        elif name is None:
            raise RibError(f"line {line}: parameter before any request")
        elif kind == "[":
            if arr is not None:
                raise RibError(f"line {line}: nested array")
            arr = []
        elif kind == "]":
            if arr is None:
                raise RibError(f"line {line}: unmatched ]")
            args.append(arr)
            arr = None
        elif arr is not None:
            arr.append(val)
        else:
            args.append(val)
    if arr is not None:
        raise RibError("unterminated array")
    if name is not None:
        yield line0, name, args


def _flat(args):
    out = []
    for a in args:
        out.extend(a if isinstance(a, list) else [a])
    return out


def _params(args):
    """Token/value pairs after the positional arguments."""
    out = {}
    i = 0
    while i < len(args):
        if isinstance(args[i], str) and i + 1 < len(args):
            v = args[i + 1]
            out[args[i]] = v if isinstance(v, list) else [v]
            i += 2
        else:
            i += 1
    return out


def _rotate(angle, x, y, z):
    n = math.sqrt(x * x + y * y + z * z)
    if n == 0.0:
        return np.eye(4)
    x, y, z = x / n, y / n, z / n
    c, s = math.cos(math.radians(angle)), math.sin(math.radians(angle))
    t = 1.0 - c
    # row-vector form: p' = p R
    return np.array([[t * x * x + c, t * x * y + s * z, t * x * z - s * y, 0.0],
                     [t * x * y - s * z, t * y * y + c, t * y * z + s * x, 0.0],
                     [t * x * z + s * y, t * y * z - s * x, t * z * z + c, 0.0],
                     [0.0, 0.0, 0.0, 1.0]])


def parse(text):
    """RIB text -> (rh_options, [rh_prim]) for the hider."""
    opts = default_options()
    prims = []
    ctm = np.eye(4)
    camera = np.eye(4)
    in_world = False
    attr = {"Cs": (1.0, 1.0, 1.0), "Os": (1.0, 1.0, 1.0), "rate": 1.0, "surface": "constant"}
    astack, tstack = [], []
    for line, req, args in _requests(text):
        flat = [a for a in _flat(args) if not isinstance(a, str)]
        try:
            if req == "Format":
                opts.xres, opts.yres = int(flat[0]), int(flat[1])
                opts.pixel_aspect = flat[2] if len(flat) > 2 and flat[2] > 0 else 1.0
            elif req == "PixelSamples":
                # RiPixelSamples truncates to integers (ri.cpp:852-856)
                opts.xsamples, opts.ysamples = int(flat[0]), int(flat[1])
            elif req == "PixelFilter":
                name = args[0]
                if name not in FILTERS:
                    raise RibError(f"line {line}: unknown pixel filter {name!r}")
                opts.filter = FILTERS[name]
                opts.filter_xwidth, opts.filter_ywidth = flat[0], flat[1]
            elif req == "ShadingRate":
                attr["rate"] = flat[0]
            elif req == "Exposure":
                opts.exposure_gain, opts.exposure_gamma = flat[0], flat[1]
            elif req == "Quantize":
                if args[0] != "rgba":
                    raise RibError(f"line {line}: only Quantize \"rgba\" is supported")
                opts.quantize_one, opts.quantize_min, opts.quantize_max, opts.quantize_dither = flat[:4]
            elif req == "Projection":
                if args[0] != "perspective":
                    raise RibError(f"line {line}: only the perspective projection is supported")
                p = _params(args[1:])
                if "fov" in p:
                    opts.fov = p["fov"][0]
            elif req == "Clipping":
                opts.clip_near, opts.clip_far = flat[0], flat[1]
            elif req == "CropWindow":
                opts.crop[:] = flat[:4]
            elif req == "Shutter":
                opts.shutter_open, opts.shutter_close = flat[0], flat[1]
            elif req == "Option":
                p = _params(args[1:])
                if args[0] == "limits":
                    if "bucketsize" in p:
                        opts.bucket_x, opts.bucket_y = int(p["bucketsize"][0]), int(p["bucketsize"][1])
                    if "eyesplits" in p:
                        opts.eyesplits = int(p["eyesplits"][0])
                    if "zthreshold" in p:
                        opts.zthreshold[:] = p["zthreshold"][:3]
            elif req == "Hider":
                p = _params(args[1:])
                if "jitter" in p:
                    opts.jitter = int(p["jitter"][0])
            elif req in ("Display", "FrameBegin", "FrameEnd", "Declare", "Sides"):
                pass
            elif req == "WorldBegin":
                camera = ctm.copy()
                ctm = np.eye(4)
                in_world = True
            elif req == "WorldEnd":
                in_world = False
            elif req == "AttributeBegin":
                astack.append(dict(attr))
                tstack.append(ctm.copy())
            elif req == "AttributeEnd":
                attr = astack.pop()
                ctm = tstack.pop()
            elif req == "TransformBegin":
                tstack.append(ctm.copy())
            elif req == "TransformEnd":
                ctm = tstack.pop()
            elif req == "Identity":
                ctm = np.eye(4)
            elif req == "Transform":
                ctm = np.array(flat[:16], dtype=np.float64).reshape(4, 4)
            elif req == "ConcatTransform":
                ctm = np.array(flat[:16], dtype=np.float64).reshape(4, 4) @ ctm
            elif req == "Translate":
                t = np.eye(4)
                t[3, :3] = flat[:3]
                ctm = t @ ctm
            elif req == "Scale":
                ctm = np.diag([flat[0], flat[1], flat[2], 1.0]) @ ctm
            elif req == "Rotate":
                ctm = _rotate(*flat[:4]) @ ctm
            elif req == "Color":
                attr["Cs"] = tuple(flat[:3])
            elif req == "Opacity":
                attr["Os"] = tuple(flat[:3])
            elif req == "Surface":
                if args[0] != "constant":
                    raise RibError(f"line {line}: only the constant surface shader is supported")
                attr["surface"] = args[0]
            elif req == "Sphere":
                if not in_world:
                    raise RibError(f"line {line}: Sphere outside WorldBegin/WorldEnd")
                m = ctm @ camera
                p = scenes.sphere(flat[0], (0.0, 0.0, 0.0), attr["Cs"], attr["Os"], attr["rate"])
                p.sphere[:] = [float(v) for v in flat[:4]]
                p.xform[0][:] = [float(v) for v in m.reshape(-1)]
                p.xform[1][:] = [float(v) for v in m.reshape(-1)]
                prims.append(p)
            elif req == "Patch":
                if not in_world:
                    raise RibError(f"line {line}: Patch outside WorldBegin/WorldEnd")
                kind = args[0]
                pts = _params(args[1:]).get("P")
                n = {"bicubic": 16, "bilinear": 4}.get(kind)
                if n is None:
                    raise RibError(f"line {line}: Patch type {kind!r} is not supported")
                if pts is None or len(pts) != 3 * n:
                    raise RibError(f"line {line}: Patch \"{kind}\" needs {3 * n} values of \"P\"")
                P = np.array(pts, dtype=np.float64).reshape(n, 3)
                m = ctm @ camera
                if not np.array_equal(m, np.eye(4)):
                    P = np.hstack([P, np.ones((n, 1))]) @ m
                    P = P[:, :3] / P[:, 3:4]
                make = scenes.bicubic if kind == "bicubic" else scenes.bilinear
                prims.append(make(P.astype(np.float32), attr["Cs"], attr["Os"], attr["rate"]))
            else:
                raise RibError(f"line {line}: request {req} is not supported by the hider front end")
        except (IndexError, TypeError) as e:
            raise RibError(f"line {line}: malformed {req}: {e}") from None
    return opts, prims


def read(path):
    with open(path) as f:
        return parse(f.read())
