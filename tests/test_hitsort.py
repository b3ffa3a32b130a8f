"""The sampling kernel's per-sample depth sort (aqsis_amd/csrc/hitsort.h)
must order hits exactly as libstdc++'s std::sort does in the reference's
Combine (imagepixel.cpp:167), ties included, since std::sort is not stable.
The same header is compiled for the host here and compared with std::sort on
random depth lists with many ties, of every length up to the device limit;
its heap-sort fallback is compared with std::partial_sort(first, last, last),
which is what std::sort falls back to at its depth limit."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROG = r'''
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>
#include "hitsort.h"
struct H { float d; uint32_t i; };
int main() {
    std::mt19937 rng(12345);
    long bad = 0, cases = 0;
    for (int rep = 0; rep < 400; ++rep)
        for (int n = 0; n <= 64; ++n) {
            const int levels = 1 + (int)(rng() % 12);     // few distinct depths: many ties
            std::vector<H> v(n);
            float d[64]; uint32_t x[64];
            for (int k = 0; k < n; ++k) {
                v[k].d = (rep % 3 == 0) ? (float)(rng() % levels) : (float)(rng() % 1000) * 0.25f;
                if (rep % 7 == 0) v[k].d = (float)(n - k);          // descending runs
                if (rep % 11 == 0) v[k].d = (float)k;                // ascending runs
                v[k].i = (uint32_t)k;
                d[k] = v[k].d; x[k] = v[k].i;
            }
            // The introsort's depth-limit fallback is std::partial_sort(f, l, l).
            std::vector<H> w = v;
            float hd[64]; uint32_t hx[64];
            for (int k = 0; k < n; ++k) { hd[k] = d[k]; hx[k] = x[k]; }
            std::partial_sort(w.begin(), w.end(), w.end(), [](const H& a, const H& b) { return a.d < b.d; });
            rhsort::heap_sort(hd, hx, 0, n);
            std::sort(v.begin(), v.end(), [](const H& a, const H& b) { return a.d < b.d; });
            rhsort::sort(d, x, n);
            ++cases;
            for (int k = 0; k < n; ++k)
                if (v[k].i != x[k] || v[k].d != d[k] || w[k].i != hx[k]) { ++bad; break; }
        }
    std::printf("%ld %ld\n", cases, bad);
    return 0;
}
'''


def test_hit_sort_matches_std_sort(tmp_path):
    src = tmp_path / "sortcheck.cpp"
    src.write_text(PROG)
    exe = tmp_path / "sortcheck"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "aqsis_amd", "csrc"), "-o", str(exe),
                    str(src)], check=True)
    cases, bad = map(int, subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split())
    assert cases == 400 * 65
    assert bad == 0


LONG_PROG = r"""
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>
#include "hitsort.h"
struct H { float d; uint32_t i; };
int main() {
    std::mt19937 rng(777);
    long bad = 0, cases = 0;
    std::vector<int> stk(2 * 64);
    const int lens[] = {0, 1, 17, 63, 64, 65, 100, 129, 255, 256, 257, 1000, 4099, 20000};
    for (int rep = 0; rep < 30; ++rep)
        for (int n : lens) {
            const int levels = 1 + (int)(rng() % 40);
            std::vector<H> v(n);
            std::vector<float> d(n);
            std::vector<uint32_t> x(n);
            for (int k = 0; k < n; ++k) {
                v[k].d = (rep % 3 == 0) ? (float)(rng() % levels) : (float)(rng() % 100000) * 0.125f;
                if (rep % 5 == 0) v[k].d = (float)(n - k);
                if (rep % 7 == 0) v[k].d = (float)((k * 7919) % 13);   // adversarial runs of ties
                v[k].i = (uint32_t)k;
                d[k] = v[k].d; x[k] = v[k].i;
            }
            std::sort(v.begin(), v.end(), [](const H& a, const H& b) { return a.d < b.d; });
            rhsort::sort_long(d.data(), x.data(), n, stk.data());
            ++cases;
            for (int k = 0; k < n; ++k)
                if (v[k].i != x[k] || v[k].d != d[k]) { ++bad; break; }
        }
    std::printf("%ld %ld\n", cases, bad);
    return 0;
}
"""


def test_long_hit_sort_matches_std_sort(tmp_path):
    """The long-list form (samples with more hits than the combine kernel
    holds in LDS) against std::sort on lists of up to 20,000 hits."""
    src = tmp_path / "sortlong.cpp"
    src.write_text(LONG_PROG)
    exe = tmp_path / "sortlong"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "aqsis_amd", "csrc"), "-o", str(exe),
                    str(src)], check=True)
    cases, bad = map(int, subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split())
    assert cases == 30 * 14
    assert bad == 0
