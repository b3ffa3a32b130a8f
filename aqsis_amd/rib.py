"""A RIB front end for the hider: the subset of the RenderMan Interface
Bytestream that drives the hot path, read into rh_options and camera-space
rh_prim records.

Requests (libs/core/api/ri.cpp; parsing follows libs/riutil/ribparser.cpp's
token rules: strings in double quotes, numbers, [ ] arrays, # comments):

  Format, PixelSamples, PixelFilter, ShadingRate, Exposure, Quantize "rgba" / "z",
  Projection "perspective" "fov", Clipping, CropWindow, Shutter,
  DepthOfField (no arguments: the pinhole camera), GeometricApproximation
  "focusfactor", ShadingInterpolation "constant" / "smooth", Attribute "cull"
  "hidden", Matte,
  Option "limits" "bucketsize" / "eyesplits" / "zthreshold",
  Hider "hidden" "jitter" / "depthfilter", Display, FrameBegin/End,
  WorldBegin/End,
  AttributeBegin/End, TransformBegin/End, Identity, Transform,
  ConcatTransform, Translate, Rotate, Scale, Color, Opacity,
  Surface "constant", Sphere, Patch "bicubic" / "bilinear" "P",
  MotionBegin [t0 t1] / MotionEnd around two requests of one kind: two
  transformations (the camera's before WorldBegin, or an object's) or two
  patches of the same type (a deforming surface).

Transformations use RenderMan's row-vector convention: a request's matrix
M is applied before the current one (CTM <- M * CTM); transformations before
WorldBegin make the camera transformation, and surfaces are posted in camera
space.  A surface posted while the camera or its own transformation moves,
or given as two motion keys, is a two-key deforming surface (nkeys = 2):
patches carry key 1's camera-space points in P1, spheres key 1's
object->camera matrix in xform[1].  Anything else raises RibError naming the
request and line.

Display requests are counted into num_displays, since each display request
draws one dither value per pixel from the shared random stream
(ddmanager.cpp:1050).  As RiDisplay does (libs/core/api/ri.cpp:995-1004), a
request whose name starts with '+' adds a display and any other replaces the
requests made so far; the name is kept without its '+'.  A mode that reads
"z" after an optional "rgb" and "a" (ri.cpp:947-964) puts z in the display
mode (display_z): added to it by a '+' request, replacing it otherwise.  Each
request dict also carries its mode bits and its resolved Quantize settings
(rh_read_display_request's arguments).  render_file writes
the first display as a TIFF file when its type is "file" or "tiff"
(tools/displays/file/display.cpp).
"""
import math
import re
import warnings

import numpy as np

from .abi import DEPTH_FILTERS, FILTERS, FLT_MAX, RH_DEPTH_MIN, default_options
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


_TRANSFORMS = ("Identity", "Transform", "ConcatTransform", "Translate", "Scale", "Rotate")


def _transform(req, flat, ctm):
    """The current transformation after one transformation request."""
    if req == "Identity":
        return np.eye(4)
    if req == "Transform":
        return np.array(flat[:16], dtype=np.float64).reshape(4, 4)
    if req == "ConcatTransform":
        return np.array(flat[:16], dtype=np.float64).reshape(4, 4) @ ctm
    if req == "Translate":
        t = np.eye(4)
        t[3, :3] = flat[:3]
        return t @ ctm
    if req == "Scale":
        return np.diag([flat[0], flat[1], flat[2], 1.0]) @ ctm
    return _rotate(*flat[:4]) @ ctm


def _patch_points(line, args, m):
    """(kind, camera-space control points) of a Patch request under m."""
    kind = args[0]
    pts = _params(args[1:]).get("P")
    n = {"bicubic": 16, "bilinear": 4}.get(kind)
    if n is None:
        raise RibError(f"line {line}: Patch type {kind!r} is not supported")
    if pts is None or len(pts) != 3 * n:
        raise RibError(f"line {line}: Patch \"{kind}\" needs {3 * n} values of \"P\"")
    P = np.array(pts, dtype=np.float64).reshape(n, 3)
    if not np.array_equal(m, np.eye(4)):
        P = np.hstack([P, np.ones((n, 1))]) @ m
        P = P[:, :3] / P[:, 3:4]
    return kind, P.astype(np.float32)


def _set_keys(prim, times, P1=None, xform1=None):
    prim.nkeys = 2
    prim.key_time[:] = [float(times[0]), float(times[1])]
    if P1 is not None:
        for i in range(len(P1)):
            prim.P1[i][:] = [float(v) for v in P1[i]]
    if xform1 is not None:
        prim.xform[1][:] = [float(v) for v in xform1.reshape(-1)]


def _apply_attrs(p, attr):
    # the surface attributes the hider reads besides colour, opacity and rate
    p.focus_factor = attr["focus"]
    p.shading_interpolation = attr["interp"]
    p.no_cull_hidden = 0 if attr["cull_hidden"] else 1
    p.matte = attr["matte"]


def _mode_has_z(mode):
    """DMode_Z of a display mode string: "rgb", then "a", then "z", each
    optional and matched as a prefix in that order (ri.cpp:947-964)."""
    i = 3 if mode.startswith("rgb") else 0
    if mode.startswith("a", i):
        i += 1
    return mode.startswith("z", i)


def parse(text, displays=None):
    """RIB text -> (rh_options, [rh_prim]) for the hider.  The Display
    requests are appended to `displays` when a list is given, as dicts with
    the name, type, mode, parameters and line."""
    opts = default_options()
    requests = []          # the display requests in force
    prims = []
    # The current transformation: one matrix, or one per motion key when it
    # moves (`keys`, with the key times in `ktimes`); likewise the camera.
    ctm, keys = np.eye(4), None
    camera, camkeys = np.eye(4), None
    ktimes = None
    in_world = False
    attr = {"Cs": (1.0, 1.0, 1.0), "Os": (1.0, 1.0, 1.0), "rate": 1.0, "surface": "constant", "focus": 1.0,
            "interp": 0, "cull_hidden": 1, "matte": 0}
    astack, tstack = [], []
    motion = None          # the open MotionBegin block: times, line, requests
    zquant = [0.0, 0.0, 0.0, 0.0]   # Quantize "z" (options.cpp:307)

    def key_matrices():
        """Object->camera matrices per key, or None when nothing moves."""
        if keys is None and camkeys is None:
            return None
        k = keys or [ctm, ctm]
        c = camkeys or [camera, camera]
        return [k[0] @ c[0], k[1] @ c[1]]

    for line, req, args in _requests(text):
        flat = [a for a in _flat(args) if not isinstance(a, str)]
        if motion is not None and req != "MotionEnd":
            if req == "MotionBegin":
                raise RibError(f"line {line}: MotionBegin inside a motion block")
            motion["reqs"].append((line, req, args, flat))
            continue
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
                # RiQuantize (ri.cpp:902-928): "rgba" the colour settings,
                # "z" the depth settings display requests with z take
                if args[0] == "rgba":
                    opts.quantize_one, opts.quantize_min, opts.quantize_max, opts.quantize_dither = flat[:4]
                elif args[0] == "z":
                    zquant = [float(v) for v in flat[:4]]
                else:
                    raise RibError(f"line {line}: only Quantize \"rgba\" and \"z\" are supported")
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
            elif req == "DepthOfField":
                # without arguments: back to the pinhole camera (ribparser.cpp:406-420)
                opts.dof[:] = [FLT_MAX, FLT_MAX, FLT_MAX] if not flat else [float(v) for v in flat[:3]]
                if flat and len(flat) != 3:
                    raise RibError(f"line {line}: DepthOfField needs fstop, focal length and focal distance")
            elif req == "ShadingInterpolation":
                # RiShadingInterpolation (ri.cpp:1406-1416); smooth by default
                if args[0] not in ("constant", "smooth"):
                    raise RibError(f"line {line}: ShadingInterpolation must be \"constant\" or \"smooth\"")
                attr["interp"] = 1 if args[0] == "constant" else 0
            elif req == "Attribute":
                p = _params(args[1:])
                if args[0] != "cull" or set(p) - {"hidden"}:
                    raise RibError(f"line {line}: only Attribute \"cull\" \"hidden\" is supported")
                attr["cull_hidden"] = int(p["hidden"][0])
            elif req == "Matte":
                # RiMatte's RtBoolean as CacheGridInfo reads it: 0 none, 2
                # matte alpha, any other value matte (micropolygon.cpp:49-55)
                v = int(flat[0])
                attr["matte"] = 0 if v == 0 else (2 if v == 2 else 1)
            elif req == "GeometricApproximation":
                if args[0] != "focusfactor":
                    raise RibError(f"line {line}: only GeometricApproximation \"focusfactor\" is supported")
                attr["focus"] = float(flat[0])
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
                if "depthfilter" in p:
                    # an unknown name leaves "min", with a warning (optioncache.cpp:96-110)
                    name = p["depthfilter"][0]
                    if name not in DEPTH_FILTERS:
                        warnings.warn(f"line {line}: invalid depthfilter {name!r}, depthfilter set to \"min\"")
                    opts.depth_filter = DEPTH_FILTERS.get(name, RH_DEPTH_MIN)
            elif req == "Display":
                if len(args) < 3 or not all(isinstance(a, str) for a in args[:3]):
                    raise RibError(f"line {line}: Display needs a name, a type and a mode")
                name = args[0]
                z = _mode_has_z(args[2])
                if name.startswith("+"):
                    name = name[1:]
                    opts.display_z = 1 if (opts.display_z or z) else 0
                else:
                    requests = []          # ClearDisplayRequests
                    opts.display_z = 1 if z else 0
                requests.append({"name": name, "type": args[1], "mode": args[2],
                                 "params": _params(args[3:]), "line": line})
            elif req in ("FrameBegin", "FrameEnd", "Declare", "Sides"):
                pass
            elif req == "MotionBegin":
                if len(flat) != 2 or not flat[0] < flat[1]:
                    raise RibError(f"line {line}: MotionBegin needs two increasing times (two motion keys)")
                motion = {"times": (flat[0], flat[1]), "line": line, "reqs": []}
            elif req == "MotionEnd":
                if motion is None:
                    raise RibError(f"line {line}: MotionEnd without MotionBegin")
                block, motion = motion, None
                reqs = block["reqs"]
                names = {r[1] for r in reqs}
                if len(reqs) != 2 or len(names) != 1:
                    raise RibError(f"line {block['line']}: a motion block needs one request per key, "
                                   f"both of one kind")
                if ktimes is not None and tuple(ktimes) != tuple(block["times"]):
                    raise RibError(f"line {block['line']}: motion blocks with different key times")
                kind = reqs[0][1]
                if kind in _TRANSFORMS:
                    base = keys or [ctm, ctm]
                    keys = [_transform(kind, reqs[k][3], base[k]) for k in range(2)]
                    ctm = keys[0]
                    ktimes = block["times"]
                elif kind == "Patch":
                    if not in_world:
                        raise RibError(f"line {block['line']}: Patch outside WorldBegin/WorldEnd")
                    mk = key_matrices() or [ctm @ camera] * 2
                    k0, P0 = _patch_points(reqs[0][0], reqs[0][2], mk[0])
                    k1, P1 = _patch_points(reqs[1][0], reqs[1][2], mk[1])
                    if k0 != k1:
                        raise RibError(f"line {block['line']}: the motion keys are patches of different types")
                    make = scenes.bicubic if k0 == "bicubic" else scenes.bilinear
                    p = make(P0, attr["Cs"], attr["Os"], attr["rate"])
                    _apply_attrs(p, attr)
                    _set_keys(p, block["times"], P1=P1)
                    prims.append(p)
                else:
                    raise RibError(f"line {block['line']}: {kind} is not supported inside a motion block")
            elif req == "WorldBegin":
                camera, camkeys = ctm.copy(), keys
                ctm, keys = np.eye(4), None
                in_world = True
            elif req == "WorldEnd":
                in_world = False
            elif req == "AttributeBegin":
                astack.append(dict(attr))
                tstack.append((ctm.copy(), keys))
            elif req == "AttributeEnd":
                attr = astack.pop()
                ctm, keys = tstack.pop()
            elif req == "TransformBegin":
                tstack.append((ctm.copy(), keys))
            elif req == "TransformEnd":
                ctm, keys = tstack.pop()
            elif req in _TRANSFORMS:
                # a static request applies to every key of a moving transformation
                ctm = _transform(req, flat, ctm)
                if keys is not None:
                    keys = [_transform(req, flat, k) for k in keys]
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
                mk = key_matrices()
                m = ctm @ camera if mk is None else mk[0]
                p = scenes.sphere(flat[0], (0.0, 0.0, 0.0), attr["Cs"], attr["Os"], attr["rate"])
                _apply_attrs(p, attr)
                p.sphere[:] = [float(v) for v in flat[:4]]
                p.xform[0][:] = [float(v) for v in m.reshape(-1)]
                p.xform[1][:] = [float(v) for v in m.reshape(-1)]
                if mk is not None:
                    _set_keys(p, ktimes, xform1=mk[1])
                prims.append(p)
            elif req == "Patch":
                if not in_world:
                    raise RibError(f"line {line}: Patch outside WorldBegin/WorldEnd")
                mk = key_matrices()
                kind, P = _patch_points(line, args, ctm @ camera if mk is None else mk[0])
                make = scenes.bicubic if kind == "bicubic" else scenes.bilinear
                p = make(P, attr["Cs"], attr["Os"], attr["rate"])
                _apply_attrs(p, attr)
                if mk is not None:
                    _set_keys(p, ktimes, P1=_patch_points(line, args, mk[1])[1])
                prims.append(p)
            else:
                raise RibError(f"line {line}: request {req} is not supported by the hider front end")
        except (IndexError, TypeError) as e:
            raise RibError(f"line {line}: malformed {req}: {e}") from None
    if motion is not None:
        raise RibError(f"line {motion['line']}: MotionBegin without MotionEnd")
    if requests:
        opts.num_displays = len(requests)
    # each request's Quantize settings (ddmanager.cpp:429-455, 848-870): its
    # own "quantize" [zero one min max] and "dither", else Quantize "z" when
    # its mode has z, else Quantize "rgba"
    from .hider import MODE_Z, mode_bits
    rgba = [opts.quantize_one, opts.quantize_min, opts.quantize_max, opts.quantize_dither]
    for r in requests:
        bits = mode_bits(r["mode"])
        q = list(zquant if bits & MODE_Z else rgba)
        pq, pd = r["params"].get("quantize"), r["params"].get("dither")
        if pq is not None:
            if len(pq) != 4 or pq[0] != 0:
                raise RibError(f"line {r['line']}: Display \"quantize\" must be [0 one min max]")
            q[0:3] = [float(v) for v in pq[1:4]]
        if pd is not None:
            q[3] = float(pd[0])
        r["mode_bits"] = bits
        r["quantize"] = tuple(q)
    if displays is not None:
        displays.extend(requests)
    return opts, prims


def read(path, displays=None):
    with open(path) as f:
        return parse(f.read(), displays)

def render_file(path, out_dir=None):
    """Render a RIB file and write its first display, when that is a "file"
    or "tiff" display of mode "rgb" or "rgba", as a TIFF file (the name
    taken relative to `out_dir`, default the RIB's folder; the "compression"
    parameter "lzw", the driver's default, or "none").  Other display types
    are rendered and not written.  Returns the paths written."""
    import os
    from .hider import Hider
    displays = []
    opts, prims = read(path, displays)
    out_dir = os.path.dirname(os.path.abspath(path)) if out_dir is None else out_dir
    files = [d for d in displays if d["type"] in ("file", "tiff")]
    for d in files:
        if d is not displays[0]:
            raise RibError(f"line {d['line']}: only the first display request can be written to a file")
        if d["mode"] not in ("rgb", "rgba"):
            raise RibError(f"line {d['line']}: file display mode {d['mode']!r} is not rgb or rgba")
        comp = d["params"].get("compression", ["lzw"])[0]
        if comp not in ("lzw", "none"):
            raise RibError(f"line {d['line']}: file display compression {comp!r} is not lzw or none")
    written = []
    with Hider(opts) as h:
        h.post_surfaces(prims)
        h.render()
        for d in files:
            target = os.path.join(out_dir, d["name"])
            h.write_tiff(target, d["params"].get("compression", ["lzw"])[0], channels=len(d["mode"]))
            written.append(target)
    return written


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="Render RIB files with the hider and write their file displays.")
    ap.add_argument("rib", nargs="+")
    ap.add_argument("-o", "--out-dir", default=None)
    a = ap.parse_args()
    for f in a.rib:
        for w in render_file(f, a.out_dir):
            print(w)
