"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the oracle): regression pins of per-sample visibility and pixels.  The
oracle must reproduce them with and without occlusion culling (CPU), and the
HIP hider must reproduce them bit for bit (GPU)."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import config_cases, frame_digest, golden_cases  # noqa: E402

CASES = golden_cases()


def _fixture(name):
    with np.load(os.path.join(HERE, "golden", f"{name}.npz")) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("cull", [False, True])
def test_oracle_reproduces_golden(oracle, name, cull):
    opts, prims = CASES[name]
    g = _fixture(name)
    r = oracle.render(opts, prims, occlusion_cull=cull)
    # Micropolygon ids number the busted micropolygons in dicing order, so
    # they depend on which occluded surfaces culling defers; these scenes
    # hold no equal-depth hits, so the hits themselves do not.
    assert np.array_equal(r["mp_id"] == 0xFFFFFFFF, g["mp_id"] == 0xFFFFFFFF)
    if cull:
        assert np.array_equal(r["mp_id"], g["mp_id"])
    assert np.array_equal(r["depth"].view(np.uint32), g["depth"].view(np.uint32))
    assert np.array_equal(r["channels"].view(np.uint32), g["channels"].view(np.uint32))
    assert np.array_equal(r["rgba8"], g["rgba8"])
    if cull:
        assert r["stats"]["spl_count"] == int(g["spl_count"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_device_reproduces_golden(name):
    from aqsis_amd.hider import Hider
    opts, prims = CASES[name]
    g = _fixture(name)
    print(f"golden {name}", flush=True)
    with Hider(opts) as h:
        h.post_surfaces(prims)
        h.render()
        ids, depth, _ = h.samples()
        assert np.array_equal(ids, g["mp_id"])
        assert np.array_equal(depth.view(np.uint32), g["depth"].view(np.uint32))
        assert float(np.max(np.abs(h.channels().astype(np.float64) - g["channels"]))) <= 1e-5
        assert np.array_equal(h.rgba8(), g["rgba8"])
        assert h.stats()["spl_count"] == int(g["spl_count"])


@pytest.mark.parametrize("name", ["hd_patches", "motion_256"])
def test_oracle_reproduces_config_digests(oracle, name):
    """The full-size configurations' digests (checked against the device by
    tests/test_configs_gpu.py) are the oracle's frames."""
    import json
    with open(os.path.join(HERE, "golden", "configs.json")) as f:
        want = json.load(f)[name]
    opts, prims = config_cases()[name]()
    r = oracle.render(opts, prims, occlusion_cull=True)
    assert frame_digest(r["mp_id"], r["depth"], r["rgbaz"], r["channels"], r["rgba8"], r["stats"]) == want
