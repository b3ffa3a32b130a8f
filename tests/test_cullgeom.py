"""Occlusion-test geometry of the host front end against the oracle's
restatement of CqOcclusionTree::canCull (libs/core/occlusion.cpp:54-196).

The product decides a cull test as "the largest depth over the rectangle of
sub-sample cells whose tree leaves meet the surface's bound lies below the
bound's minimum z" (aqsis_amd/csrc/cullgeom.h, tree boxes from rh::cull_trees
in front.cpp, the rectangle maximum read from a max pyramid).  A small C++
harness compiled here runs that code on random regions, depths and bounds,
including bounds on leaf edges, odd region sizes, 3x3 and 1x1 samples and
bounds outside the region; the oracle runs the reference's tree traversal on
the same leaves.  Every answer must agree."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "aqsis_amd", "csrc")

HARNESS = r'''
#include "front.h"
#include <cstdio>
#include <vector>
// stdin: cases of  sx0 sy0 w h xs ys  bx0 by0 bx1 by1 z  then w*xs*h*ys depths
// stdout: per case, 1 when the product's test culls, 0 otherwise
int main() {
    int sx0, sy0, w, h, xs, ys;
    float b[5];
    while (std::scanf("%d %d %d %d %d %d %f %f %f %f %f", &sx0, &sy0, &w, &h, &xs, &ys, &b[0], &b[1], &b[2], &b[3],
                      &b[4]) == 11) {
        rh::Shape s;
        s.nx = xs; s.ny = ys; s.ns = xs * ys;
        rh::Shape::BucketInfo bi;
        bi.x = sx0; bi.y = sy0; bi.w = w; bi.h = h;
        bi.sx0 = sx0; bi.sy0 = sy0; bi.sx1 = sx0 + w; bi.sy1 = sy0 + h;
        s.buckets.assign(1, bi);
        std::vector<CullTreeDev> trees;
        const uint64_t nf = rh::cull_trees(s, trees);
        const CullTreeDev& t = trees[0];
        const int cw = w * xs, ch = h * ys;
        std::vector<float> pyr(nf);
        for (int i = 0; i < cw * ch; ++i) std::scanf("%f", &pyr[i]);
        // the pyramid as the device builds it
        float* L = pyr.data();
        int lw = cw, lh = ch;
        for (int l = 1; l < t.levels; ++l) {
            float* U = L + (size_t)lw * lh;
            const int uw = cull_level_dim(cw, l), uh = cull_level_dim(ch, l);
            for (int y = 0; y < uh; ++y)
                for (int x = 0; x < uw; ++x) {
                    float m = L[(size_t)(2 * y) * lw + 2 * x];
                    if (2 * x + 1 < lw) m = m < L[(size_t)(2 * y) * lw + 2 * x + 1] ? L[(size_t)(2 * y) * lw + 2 * x + 1] : m;
                    if (2 * y + 1 < lh) {
                        m = m < L[(size_t)(2 * y + 1) * lw + 2 * x] ? L[(size_t)(2 * y + 1) * lw + 2 * x] : m;
                        if (2 * x + 1 < lw)
                            m = m < L[(size_t)(2 * y + 1) * lw + 2 * x + 1] ? L[(size_t)(2 * y + 1) * lw + 2 * x + 1] : m;
                    }
                    U[(size_t)y * uw + x] = m;
                }
            L = U; lw = uw; lh = uh;
        }
        int cx0 = 0, cx1 = -1, cy0 = 0, cy1 = -1;
        const int kind = cull_cells(t, b[0], b[1], b[2], b[3], cx0, cx1, cy0, cy1);
        float m = -1.0f;
        if (kind == 2) m = 0.0f;
        if (kind != 0 && cx0 <= cx1 && cy0 <= cy1) {
            const float pm = pyramid_max(pyr.data(), cw, ch, cx0, cx1, cy0, cy1);
            // the pyramid agrees with a direct scan of the rectangle
            float dm = -3.402823466e+38f;
            for (int y = cy0; y <= cy1; ++y)
                for (int x = cx0; x <= cx1; ++x) dm = dm < pyr[(size_t)y * cw + x] ? pyr[(size_t)y * cw + x] : dm;
            if (pm != dm) { std::printf("pyramid mismatch\n"); return 3; }
            m = m < pm ? pm : m;
        }
        std::printf("%d\n", m < b[4] ? 1 : 0);
    }
    return 0;
}
'''


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not installed")
    d = tmp_path_factory.mktemp("cullgeom")
    src = d / "h.cpp"
    src.write_text(HARNESS)
    exe = d / "h"
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-I", CSRC, str(src),
                    os.path.join(CSRC, "front.cpp"), "-o", str(exe), "-lpthread"], check=True)
    return str(exe)


def _cases(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        xs, ys = [(4, 4), (2, 2), (3, 3), (1, 1), (8, 8), (2, 3)][rng.integers(0, 6)]
        w, h = int(rng.integers(1, 20)), int(rng.integers(1, 20))
        sx0, sy0 = int(rng.integers(-2, 60)), int(rng.integers(-2, 60))
        cw, ch = w * xs, h * ys
        depths = rng.uniform(1.0, 20.0, (ch, cw)).astype(np.float32)
        # some never-hit samples
        depths[rng.uniform(size=(ch, cw)) < 0.1] = np.float32(3.4028234663852886e38)
        # bounds: random, on leaf edges (multiples of 1/xs), or outside
        k = rng.integers(0, 3)
        if k == 0:
            x0 = sx0 + rng.uniform(-3, w + 1)
            y0 = sy0 + rng.uniform(-3, h + 1)
            x1 = x0 + rng.uniform(0, w + 2)
            y1 = y0 + rng.uniform(0, h + 2)
        elif k == 1:
            x0 = sx0 + rng.integers(-2, w * xs + 2) / xs
            y0 = sy0 + rng.integers(-2, h * ys + 2) / ys
            x1 = x0 + rng.integers(0, w * xs + 2) / xs
            y1 = y0 + rng.integers(0, h * ys + 2) / ys
        else:
            x0 = sx0 + w + rng.uniform(0.0, 40.0)
            y0 = sy0 + rng.uniform(-5, h + 5)
            x1, y1 = x0 + 3.0, y0 + 3.0
        bound = np.array([x0, y0, x1, y1], np.float32)
        # z: above, below or exactly at the region's depths
        m = float(depths.max())
        z = np.float32([rng.uniform(0.5, 25.0), m, np.nextafter(np.float32(m), np.float32(0)),
                        float(depths.min())][rng.integers(0, 4)])
        out.append((sx0, sy0, w, h, xs, ys, bound, z, depths))
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_cull_test_matches_occlusion_tree(oracle, harness, seed):
    import ctypes as C
    cases = _cases(seed, 400)
    lines = []
    for sx0, sy0, w, h, xs, ys, bound, z, depths in cases:
        lines.append(f"{sx0} {sy0} {w} {h} {xs} {ys} " + " ".join(repr(float(v)) for v in bound) +
                     f" {float(z)!r}\n" + " ".join(repr(float(v)) for v in depths.reshape(-1)) + "\n")
    res = subprocess.run([harness], input="".join(lines), capture_output=True, text=True, check=True)
    got = [int(v) for v in res.stdout.split()]
    assert len(got) == len(cases)
    culls = 0
    for (sx0, sy0, w, h, xs, ys, bound, z, depths), g in zip(cases, got):
        b6 = np.array([bound[0], bound[1], bound[2], bound[3], z, z + 1], np.float32)
        want = oracle.lib().oracle_can_cull(sx0, sy0, w, h, xs, ys, depths.ctypes.data_as(C.c_void_p),
                                            b6.ctypes.data_as(C.c_void_p))
        assert g == want, (sx0, sy0, w, h, xs, ys, bound.tolist(), float(z))
        culls += want
    # both outcomes are exercised
    assert 0 < culls < len(cases)
