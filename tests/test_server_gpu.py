"""Serving loop: jobs of mixed frame shapes, sample counts and opacity run
concurrently on one GPU; each result is identical to rendering the job
alone, also when a tight memory budget forces queueing and eviction."""
import numpy as np
import pytest

from aqsis_amd import scenes
from parity_cases import CASES

pytestmark = pytest.mark.gpu


def _jobs():
    return [
        CASES["spheres_config"](),
        CASES["patches_small"](),
        CASES["patches_transparent"](),
        CASES["spheres_box_2x2"](),
        scenes.patches_scene(seed=41, n=200, res=(128, 128), samples=(8, 8), opacity=(0.2, 1.0),
                             filt=("gaussian", 2.0, 2.0)),
        CASES["patches_small"](),          # same shape again: a resident context is reused
        scenes.spheres_scene(seed=43, n=20, res=(96, 64), samples=(2, 2)),
    ]


def _alone(opts, prims):
    from aqsis_amd.hider import Hider
    with Hider(opts) as h:
        h.post_surfaces(prims)
        h.render()
        return h.channels(), h.rgba8()


@pytest.mark.parametrize("slots,budget", [(4, 64 << 30), (3, 800 << 20)])
def test_results_same_alone_and_in_the_mix(slots, budget):
    from aqsis_amd.hider import Server
    jobs = _jobs()
    print(f"server slots={slots} budget={budget}", flush=True)
    ref = [_alone(o, p) for o, p in jobs]
    with Server(slots=slots, memory_budget=budget) as srv:
        ids = [srv.submit(o, p) for o, p in jobs]
        for k, jid in enumerate(ids):
            info = srv.wait(jid)
            assert info["status"] == 0 and info["state"] == 2
            ch, rgba = srv.result(jid, info["width"], info["height"])
            assert np.array_equal(ch.view(np.uint32), ref[k][0].view(np.uint32)), k
            assert np.array_equal(rgba, ref[k][1]), k
            srv.release(jid)
        resident, high = srv.memory()
        assert resident <= high and resident <= budget


def test_job_over_budget_is_refused():
    from aqsis_amd.hider import HiderError, Server
    opts, prims = CASES["patches_small"]()
    with Server(slots=1, memory_budget=1 << 20) as srv:
        jid = srv.submit(opts, prims)          # returns at once
        with pytest.raises(HiderError, match="budget"):
            srv.wait(jid)
