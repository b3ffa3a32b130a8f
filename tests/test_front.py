"""CPU checks of the product library: it loads, exports every entry point
include/aqsis_hider.h declares, validates its inputs, and its host front end
(sample plan, split/dice schedule) agrees with the oracle.  No GPU needed."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from aqsis_amd import scenes
from aqsis_amd.abi import RhOptions, RhPrim, RhStats, default_options
from parity_cases import CASES

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_entry_points():
    text = open(os.path.join(ROOT, "include", "aqsis_hider.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    from aqsis_amd import hider
    lib = hider.lib()
    names = _declared_entry_points()
    assert "rh_create" in names and "rh_render" in names
    for name in names:
        assert hasattr(lib, name), name


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors have the C compiler's sizes and field offsets."""
    import subprocess
    from aqsis_amd.hider import RhFootprint, RhTuning, RhServerOptions, RhJobInfo
    src = tmp_path / "layout.c"
    fields = {"rh_options": [f for f, _ in RhOptions._fields_], "rh_prim": [f for f, _ in RhPrim._fields_],
              "rh_stats": [f for f, _ in RhStats._fields_], "rh_footprint": [f for f, _ in RhFootprint._fields_],
              "rh_tuning": [f for f, _ in RhTuning._fields_],
              "rh_server_options": [f for f, _ in RhServerOptions._fields_],
              "rh_job_info": [f for f, _ in RhJobInfo._fields_]}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{ROOT}/include/aqsis_hider.h"',
             'int main(void) {']
    for st, fs in fields.items():
        lines.append(f'printf("%zu\\n", sizeof({st}));')
        for f in fs:
            lines.append(f'printf("%zu\\n", offsetof({st}, {f}));')
    lines.append("return 0; }")
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = []
    for cls in (RhOptions, RhPrim, RhStats, RhFootprint, RhTuning, RhServerOptions, RhJobInfo):
        want.append(C.sizeof(cls))
        want += [getattr(cls, f).offset for f, _ in cls._fields_]
    assert got == want


def test_create_without_gpu_fails_loudly_or_works():
    from aqsis_amd import hider
    opts, prims = scenes.spheres_scene(n=2, res=(32, 32))
    try:
        import torch  # noqa: F401
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(hider.HiderError, match="no HIP device"):
        hider.Hider(opts)
    with pytest.raises(hider.HiderError, match="no HIP device"):
        hider.Server(slots=2)


@pytest.mark.parametrize("field,value,msg", [
    ("xres", 0, "resolution"), ("xsamples", 0, "PixelSamples"), ("filter_xwidth", 0.0, "filter widths"),
    ("fov", 200.0, "fov"), ("bucket_x", 2, "bucket size"), ("filter", 42, "unknown pixel filter"),
])
def test_invalid_options_are_rejected(field, value, msg):
    from aqsis_amd import hider
    opts, prims = scenes.spheres_scene(n=2, res=(32, 32))
    setattr(opts, field, value)
    with pytest.raises(hider.HiderError, match=msg):
        hider.estimate(opts, prims)


def test_unsupported_surfaces_are_rejected():
    from aqsis_amd import hider
    opts, prims = scenes.spheres_scene(n=2, res=(32, 32))
    prims[1].type = 99
    with pytest.raises(hider.HiderError, match="unknown surface type"):
        hider.estimate(opts, prims)


@pytest.mark.parametrize("case", list(CASES))
def test_schedule_matches_oracle(oracle, case):
    from aqsis_amd import hider
    opts, prims = CASES[case]()
    fp = hider.estimate(opts, prims)
    st = oracle.render(opts, prims, occlusion_cull=False, samples=False)["stats"]
    assert (fp["grids"], fp["micropolygons"], fp["surfaces"]) == (st["grids"], st["micropolygons"], st["surfaces"])
    assert fp["list_capacity"] >= st["mp_bucket_refs"]


@pytest.mark.parametrize("case", list(CASES))
def test_sample_plan_matches_oracle(oracle, case):
    from aqsis_amd import hider
    opts, _ = CASES[case]()
    w, h, ew, eh, spp = oracle.frame_size(opts)
    a, da = hider.sample_positions(opts, (ew, eh, spp), (w, h))
    b, db = oracle.sample_positions(opts)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(da.view(np.uint32), db.view(np.uint32))


def test_sample_plan_at_hd(oracle):
    """The full HD frame shape: 1920x1080 at 4x4 with catmull-rom 3x3."""
    from aqsis_amd import hider
    opts = default_options(xres=1920, yres=1080, fov=45.0, xsamples=4, ysamples=4, filter="catmull-rom",
                           filter_xwidth=3.0, filter_ywidth=3.0)
    w, h, ew, eh, spp = oracle.frame_size(opts)
    a, da = hider.sample_positions(opts, (ew, eh, spp), (w, h))
    b, db = oracle.sample_positions(opts)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(da, db)


def test_scene_generators_are_deterministic():
    a = scenes.prim_array(scenes.patches_scene(seed=3, n=20, res=(64, 64))[1])
    b = scenes.prim_array(scenes.patches_scene(seed=3, n=20, res=(64, 64))[1])
    assert bytes(a) == bytes(b)
    rib = scenes.to_rib(*scenes.spheres_scene(n=3, res=(32, 32)))
    assert rib.count("Sphere ") == 3 and "WorldEnd" in rib
