"""The RIB front end (aqsis_amd/rib.py): scenes written with scenes.to_rib
read back to identical options and surface records, transforms follow
RenderMan's row-vector convention, and unsupported requests are refused."""
import ctypes as C

import numpy as np
import pytest

from aqsis_amd import rib, scenes
from aqsis_amd.abi import RhOptions, RhPrim
from parity_cases import CASES


def _same_struct(a, b):
    return bytes(a) == bytes(b)


def _static_keys_cleared(p):
    """A one-key surface with its (unused) key-1 fields zeroed."""
    q = RhPrim.from_buffer_copy(bytes(p))
    if q.nkeys != 2:
        C.memset(C.addressof(q) + RhPrim.key_time.offset, 0, C.sizeof(q.key_time))
        C.memset(C.addressof(q) + RhPrim.P1.offset, 0, C.sizeof(q.P1))
    return q


@pytest.mark.parametrize("case", ["spheres_config", "patches_crop_odd", "patches_transparent_8x8_zthreshold",
                                  "spheres_exposure_mitchell", "patches_gridsampler", "motion_transparent_8x8",
                                  "motion_mixed_opaque", "motion_spheres", "decal_dof", "spheres_dof_transparent",
                                  "patches_cull_flag_mixed", "constant_interpolation_mixed",
                                  "constant_interpolation_motion", "depth_midpoint_z", "depth_average_z",
                                  "depth_midpoint_no_z", "matte_mixed", "matte_max_z_motion"])
def test_round_trip(case):
    opts, prims = CASES[case]()
    text = scenes.to_rib(opts, prims)
    o2, p2 = rib.parse(text)
    assert _same_struct(opts, o2)
    assert len(p2) == len(prims)
    for a, b in zip(prims, p2):
        assert _same_struct(_static_keys_cleared(a), b)


def test_transforms_and_state():
    text = """
    # camera: move the world 10 units away
    Translate 0 0 10
    WorldBegin
      Color [0.5 0.25 1]
      AttributeBegin
        Translate 1 2 3
        Scale 2 2 2
        Sphere 1 -1 1 360
      AttributeEnd
      Opacity [0.5 0.5 0.5]
      Rotate 90 0 0 1
      Patch "bilinear" "P" [0 0 0  1 0 0  0 1 0  1 1 0]
    WorldEnd
    """
    opts, prims = rib.parse(text)
    s, pb = prims
    m = np.array(s.xform[0], np.float64).reshape(4, 4)
    # p = (0,0,0) of the sphere's object space -> (1,2,13) in camera space
    assert np.allclose(np.array([0, 0, 0, 1.0]) @ m, [1, 2, 13, 1])
    assert np.allclose(np.array([1, 0, 0, 1.0]) @ m, [3, 2, 13, 1])      # scaled by 2 first
    assert list(s.Cs) == [0.5, 0.25, 1.0] and list(s.Os) == [1.0, 1.0, 1.0]
    P = np.array([list(pb.P[i]) for i in range(4)])
    assert np.allclose(P[1], [0, 1, 10], atol=1e-6)                        # (1,0,0) rotated to (0,1,0)
    assert list(pb.Os) == [0.5, 0.5, 0.5] and list(pb.Cs) == [0.5, 0.25, 1.0]


def test_motion_blocks():
    """MotionBegin around two camera transformations, two object
    transformations or two patches makes two-key surfaces; static requests
    after a moving one apply to both keys."""
    text = """
    Shutter 0 1
    MotionBegin [0 1]
      Translate 0 0 10
      Translate 1 0 10
    MotionEnd
    WorldBegin
      Patch "bilinear" "P" [0 0 0  1 0 0  0 1 0  1 1 0]
      AttributeBegin
        MotionBegin [0 1]
          Translate 0 0 0
          Translate 0 2 0
        MotionEnd
        Scale 2 2 2
        Sphere 1 -1 1 360
      AttributeEnd
      MotionBegin [0 1]
        Patch "bilinear" "P" [0 0 0  1 0 0  0 1 0  1 1 0]
        Patch "bilinear" "P" [0 0 1  1 0 1  0 1 1  1 1 1]
      MotionEnd
    WorldEnd
    """
    opts, prims = rib.parse(text)
    assert (opts.shutter_open, opts.shutter_close) == (0.0, 1.0)
    pc, s, pd = prims
    # camera motion: the world moves by +1 in x between the keys
    P = np.array([list(pc.P[i]) for i in range(4)]); P1 = np.array([list(pc.P1[i]) for i in range(4)])
    assert pc.nkeys == 2 and list(pc.key_time) == [0.0, 1.0]
    assert np.allclose(P1 - P, [[1, 0, 0]] * 4) and np.allclose(P[:, 2], 10)
    # object motion with a static Scale after it, under the moving camera
    m0 = np.array(s.xform[0], np.float64).reshape(4, 4); m1 = np.array(s.xform[1], np.float64).reshape(4, 4)
    assert s.nkeys == 2
    assert np.allclose(np.array([1, 0, 0, 1.0]) @ m0, [2, 0, 10, 1])
    assert np.allclose(np.array([1, 0, 0, 1.0]) @ m1, [3, 2, 10, 1])
    # deforming patch under the moving camera: each key through its own camera matrix
    D = np.array([list(pd.P[i]) for i in range(4)]); D1 = np.array([list(pd.P1[i]) for i in range(4)])
    assert np.allclose(D1 - D, [[1, 0, 1]] * 4)


@pytest.mark.parametrize("text,msg", [
    ('Surface "plastic"', "constant"),
    ('Projection "orthographic"', "perspective"),
    ('WorldBegin\nNuPatch 4 4', "not supported"),
    ('WorldBegin\nPatch "bicubic" "P" [0 0 0]', "48 values"),
    ('Sphere 1 -1 1 360', "outside WorldBegin"),
    ('PixelFilter "lanczos" 2 2', "unknown pixel filter"),
    ('MotionBegin [0 1]\nTranslate 1 0 0\nMotionEnd', "one request per key"),
    ('MotionBegin [0 0.5 1]\nTranslate 1 0 0\nTranslate 2 0 0\nTranslate 3 0 0\nMotionEnd', "two increasing"),
    ('MotionBegin [0 1]\nTranslate 1 0 0\nScale 2 2 2\nMotionEnd', "one kind"),
    ('MotionBegin [0 1]\nTranslate 1 0 0', "without MotionEnd"),
    ('WorldBegin\nMotionBegin [0 1]\nSphere 1 -1 1 360\nSphere 2 -2 2 360\nMotionEnd', "not supported inside"),
])
def test_refused(text, msg):
    with pytest.raises(rib.RibError, match=msg):
        rib.parse(text)


def test_display_requests():
    """Display requests are collected and counted into num_displays (each
    draws one dither value per pixel from the shared stream); a '+' name adds
    a display and is recorded without the '+'."""
    text = ('Display "a.tif" "file" "rgba" "compression" "none"\n'
            'Display "+b" "framebuffer" "rgb"\n'
            'Format 32 24 1\nWorldBegin\nSphere 1 -1 1 360\nWorldEnd\n')
    displays = []
    opts, _ = rib.parse(text, displays)
    assert opts.num_displays == 2
    assert [(d["name"], d["type"], d["mode"]) for d in displays] == [("a.tif", "file", "rgba"), ("b", "framebuffer", "rgb")]
    assert displays[0]["params"] == {"compression": ["none"]}
    assert rib.parse(text.split("\n", 2)[2])[0].num_displays == 1  # none given: the default one
    with pytest.raises(rib.RibError, match="line 1: Display needs"):
        rib.parse('Display "a.tif" "file"\n')


def test_display_without_plus_replaces_earlier_requests():
    """RiDisplay with a name that has no leading '+' clears the display
    requests made before it (libs/core/api/ri.cpp:995-1004): two plain
    Display lines leave one display, the second, and one dither draw per
    pixel."""
    text = ('Display "a.tif" "file" "rgba"\n'
            'Display "b.tif" "tiff" "rgb"\n'
            'Format 32 24 1\nWorldBegin\nSphere 1 -1 1 360\nWorldEnd\n')
    displays = []
    opts, _ = rib.parse(text, displays)
    assert opts.num_displays == 1
    assert [(d["name"], d["type"], d["mode"], d["line"]) for d in displays] == [("b.tif", "tiff", "rgb", 2)]
    text3 = 'Display "a" "file" "rgba"\nDisplay "+b" "file" "rgb"\nDisplay "c" "framebuffer" "rgb"\nDisplay "+d" "file" "rgb"\n'
    displays = []
    opts, _ = rib.parse(text3, displays)
    assert opts.num_displays == 2
    assert [d["name"] for d in displays] == ["c", "d"]


@pytest.mark.parametrize("first,second,msg", [
    ('"a.tif" "file" "z"', None, "mode 'z'"),
    ('"a.tif" "tiff" "rgba" "compression" "jpeg"', None, "compression 'jpeg'"),
    ('"a" "framebuffer" "rgba"', '"+b.tif" "file" "rgba"', "only the first display"),
])
def test_file_displays_refused_before_rendering(tmp_path, first, second, msg):
    lines = [f"Display {first}"] + ([f"Display {second}"] if second else [])
    path = tmp_path / "s.rib"
    path.write_text("\n".join(lines) + "\nFormat 16 16 1\nWorldBegin\nWorldEnd\n")
    with pytest.raises(rib.RibError, match=msg):
        rib.render_file(str(path))


def test_display_request_quantize_settings():
    """Each request's Quantize settings (ddmanager.cpp:429-455, 848-870):
    its own "quantize" [0 one min max] / "dither", else Quantize "z" for a
    mode with z, else Quantize "rgba"."""
    text = """Quantize "rgba" 255 0 255 0.5
    Quantize "z" 1000 0 65535 1
    Display "a.tif" "file" "rgba"
    Display "+d.tif" "file" "z"
    Display "+e.tif" "file" "rgbaz" "quantize" [0 65535 0 65535] "dither" [0.25]
    WorldBegin
    Sphere 1 -1 1 360
    WorldEnd
    """
    displays = []
    opts, _ = rib.parse(text, displays)
    assert [d["mode_bits"] for d in displays] == [3, 4, 7]
    assert displays[0]["quantize"] == (255.0, 0.0, 255.0, 0.5)
    assert displays[1]["quantize"] == (1000.0, 0.0, 65535.0, 1.0)
    assert displays[2]["quantize"] == (65535.0, 0.0, 65535.0, 0.25)
    assert opts.num_displays == 3 and opts.display_z == 1
    with pytest.raises(rib.RibError, match="quantize"):
        rib.parse('Display "a" "file" "rgba" "quantize" [1 255 0 255]\nWorldBegin\nWorldEnd\n')
