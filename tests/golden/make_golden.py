"""Writes the golden fixtures of tests/golden/ from the oracle.

The reference hider cannot be built in this environment (it needs Boost,
flex and bison, and reaches into the renderer's singletons), so these
fixtures are regression pins of the oracle's restatement, not reference
output: per-sample winning micropolygon ids and depths, filtered channels, 8-bit
pixels and the sample-test count of small scenes, with occlusion culling on
(the reference's default, Attribute "cull" "hidden" 1).

Usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from aqsis_amd import scenes  # noqa: E402
from oracle import loader  # noqa: E402


def golden_cases():
    return {
        "spheres_64": scenes.spheres_scene(seed=2, n=16, res=(64, 64)),
        "patches_64x48": scenes.patches_scene(seed=5, n=40, res=(64, 48), samples=(2, 2),
                                              filt=("catmull-rom", 3.0, 3.0)),
    }


def main():
    for name, (opts, prims) in golden_cases().items():
        r = loader.render(opts, prims, occlusion_cull=True)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), mp_id=r["mp_id"], depth=r["depth"],
                            channels=r["channels"], rgba8=r["rgba8"],
                            spl_count=np.uint64(r["stats"]["spl_count"]))
        print(name, os.path.getsize(os.path.join(HERE, f"{name}.npz")), "bytes")


if __name__ == "__main__":
    main()
