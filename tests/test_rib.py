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


@pytest.mark.parametrize("case", ["spheres_config", "patches_crop_odd", "patches_transparent_8x8_zthreshold",
                                  "spheres_exposure_mitchell", "patches_gridsampler"])
def test_round_trip(case):
    opts, prims = CASES[case]()
    text = scenes.to_rib(opts, prims)
    o2, p2 = rib.parse(text)
    assert _same_struct(opts, o2)
    assert len(p2) == len(prims)
    for a, b in zip(prims, p2):
        assert _same_struct(a, b)


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


@pytest.mark.parametrize("text,msg", [
    ('Surface "plastic"', "constant"),
    ('Projection "orthographic"', "perspective"),
    ('WorldBegin\nNuPatch 4 4', "not supported"),
    ('WorldBegin\nPatch "bicubic" "P" [0 0 0]', "48 values"),
    ('Sphere 1 -1 1 360', "outside WorldBegin"),
    ('PixelFilter "lanczos" 2 2', "unknown pixel filter"),
])
def test_refused(text, msg):
    with pytest.raises(rib.RibError, match=msg):
        rib.parse(text)
