"""Every device layout of a frame gives the same result: the static grid,
the per-bucket launches, and the batched form (cost-split work items,
persistent workgroups, one hipGraph per frame) under each kernel variant and
split cost.  The graph is replayed several times with eager work and
synchronisation in between, and every replay is checked against the eager
frame."""
import numpy as np
import pytest

from parity_cases import CASES

pytestmark = pytest.mark.gpu


def _frame(h):
    ids, depth, rgbaz = h.samples()
    return {"ids": ids, "depth": depth.view(np.uint32), "rgbaz": rgbaz.view(np.uint32),
            "channels": h.channels().view(np.uint32), "rgba8": h.rgba8(),
            "spl": (h.stats()["spl_count"], h.stats()["spl_bound_hits"], h.stats()["spl_hits"])}


def _same(a, b):
    for k in a:
        if k == "spl":
            assert a[k] == b[k]
        else:
            assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("case", ["patches_small", "patches_transparent", "spheres_config", "patches_8x8",
                                  "motion_mixed_opaque"])
def test_layouts_give_identical_frames(case):
    from aqsis_amd.hider import Hider
    opts, prims = CASES[case]()
    print(f"layouts {case}", flush=True)
    with Hider(opts) as h:
        h.post_surfaces(prims)
        h.render()
        ref = _frame(h)
        h.set_tuning(mode="plain")
        h.render_device()
        _same(ref, _frame(h))
        for lds in (216, 96):
            for bpw in (4, 2):
                for split in (1 << 10, 1 << 16):
                    h.set_tuning(mode="batched", lds_records=lds, blocks_per_wave=bpw, split_cost=split)
                    for _ in range(3):
                        h.render_device()          # graph replay
                        _same(ref, _frame(h))
                        h.set_tuning(mode="grid")  # eager frame in between
                        h.render_device()
                        _same(ref, _frame(h))
                        h.set_tuning(mode="batched", lds_records=lds, blocks_per_wave=bpw, split_cost=split)
                    h.render_stages()              # the same launches, one by one
                    _same(ref, _frame(h))
                    st = h.stats()
                    assert st["ms_sample"] > 0 and st["ms_dice"] > 0 and st["ms_filter"] > 0


def test_autotune_keeps_results():
    from aqsis_amd.hider import Hider
    opts, prims = CASES["patches_small"]()
    with Hider(opts) as h:
        h.post_surfaces(prims)
        h.render()
        ref = _frame(h)
        t = h.autotune()
        assert t["mode"] in (0, 2) and t["frame_ms"] > 0
        h.render_device()
        _same(ref, _frame(h))
