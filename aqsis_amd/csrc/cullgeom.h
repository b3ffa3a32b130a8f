// Geometry of a bucket's occlusion tree (CqOcclusionTree, libs/core/occlusion.cpp),
// shared by the host front end (g++) and the HIP kernels (hipcc).
//
// The reference's tree covers a bucket's sample region with 2^dx x 2^dy leaf
// cells, one per sub-sample of the region (cells beyond the region are
// padding at depth 0), and splits alternately in x and y at the float
// midpoint of a node's box (occlusion.cpp:176-190).  A surface is culled when
// every leaf whose box meets its raster bound, cropped to the tree's box,
// holds a depth below the bound's minimum z (canCull, :143-196).  The leaves
// whose boxes meet a bound form a rectangle of cells: the first and last cell
// along each axis follow from the same midpoints, so a cull test is a
// maximum over a rectangle of sub-sample depths.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#define RH_CG __host__ __device__
#else
#define RH_CG
#endif

// One bucket's tree (setupTree, occlusion.cpp:54-107).
struct CullTreeDev {
    float x0, y0, x1, y1;     // the tree's box (m_treeBoundMin, m_treeBoundMax)
    int dx, dy;               // splits along x and along y
    int cw, ch;               // cells holding samples: region width * xs, height * ys
    int sx0, sy0;             // the sample region's first pixel
    uint32_t pyr;             // first float of the bucket's depth pyramid
    int levels;               // pyramid levels (level 0: cw x ch cells)
};

// Leaf cells [i0, i1] along one axis of depth `depth` over [lo, hi] whose
// boxes meet [t0, t1] (inclusive tests, as canCull's node test).  Returns
// false when none does.
RH_CG inline bool cull_leaf_range(float lo, float hi, int depth, float t0, float t1, int& i0, int& i1) {
    float a = lo, b = hi;
    int first = 0;
    for (int l = 0; l < depth; ++l) {
        const float m = (b + a) * 0.5f;   // TqFloat avgX = 0.5 * (node.maxX + node.minX)
        if (m >= t0) b = m;
        else { a = m; first += 1 << (depth - 1 - l); }
    }
    if (b < t0) return false;
    a = lo; b = hi;
    int last = 0;
    for (int l = 0; l < depth; ++l) {
        const float m = (b + a) * 0.5f;
        if (m <= t1) { a = m; last += 1 << (depth - 1 - l); }
        else b = m;
    }
    if (a > t1) return false;
    if (first > last) return false;
    i0 = first;
    i1 = last;
    return true;
}

// The cells of a cull test of raster bound (bx0, by0)-(bx1, by1) in tree t:
// the bound is cropped to the tree's box, and cells past the sampled ones are
// padding.  Returns 0 when no leaf meets the bound, 1 when the rectangle
// [cx0, cx1] x [cy0, cy1] of sampled cells is the whole answer, 2 when
// padding leaves (depth 0) meet it too (the rectangle may then be empty:
// cx0 > cx1 or cy0 > cy1).
RH_CG inline int cull_cells(const CullTreeDev& t, float bx0, float by0, float bx1, float by1, int& cx0, int& cx1,
                            int& cy0, int& cy1) {
    const float tminX = bx0 > t.x0 ? bx0 : t.x0, tminY = by0 > t.y0 ? by0 : t.y0;
    const float tmaxX = bx1 < t.x1 ? bx1 : t.x1, tmaxY = by1 < t.y1 ? by1 : t.y1;
    int i0, i1, j0, j1;
    if (!cull_leaf_range(t.x0, t.x1, t.dx, tminX, tmaxX, i0, i1)) return 0;
    if (!cull_leaf_range(t.y0, t.y1, t.dy, tminY, tmaxY, j0, j1)) return 0;
    const bool pad = i1 >= t.cw || j1 >= t.ch;
    cx0 = i0;
    cy0 = j0;
    cx1 = i1 < t.cw ? i1 : t.cw - 1;
    cy1 = j1 < t.ch ? j1 : t.ch - 1;
    return pad ? 2 : 1;
}

// Dimensions of level l of a cw x ch max pyramid (level l+1 holds the
// maximum of each 2x2 group of level l).
RH_CG inline int cull_level_dim(int n, int l) { return ((n - 1) >> l) + 1; }

// The largest value over cells [x0, x1] x [y0, y1] of a cw x ch max
// pyramid stored level after level (level 0 first, row-major): the odd
// edges of the rectangle are read at each level and the rest goes up one.
RH_CG inline float pyramid_max(const float* L, int cw, int ch, int x0, int x1, int y0, int y1) {
    float m = -3.402823466e+38f;
    int w = cw, h = ch;
    for (int l = 0; x0 <= x1 && y0 <= y1; ++l) {
        if (x0 & 1) {
            for (int y = y0; y <= y1; ++y) m = m < L[(size_t)y * w + x0] ? L[(size_t)y * w + x0] : m;
            ++x0;
        }
        if (x0 <= x1 && !(x1 & 1)) {
            for (int y = y0; y <= y1; ++y) m = m < L[(size_t)y * w + x1] ? L[(size_t)y * w + x1] : m;
            --x1;
        }
        if (x0 > x1) break;
        if (y0 & 1) {
            for (int x = x0; x <= x1; ++x) m = m < L[(size_t)y0 * w + x] ? L[(size_t)y0 * w + x] : m;
            ++y0;
        }
        if (y0 <= y1 && !(y1 & 1)) {
            for (int x = x0; x <= x1; ++x) m = m < L[(size_t)y1 * w + x] ? L[(size_t)y1 * w + x] : m;
            --y1;
        }
        if (y0 > y1) break;
        L += (size_t)w * h;
        w = cull_level_dim(cw, l + 1);
        h = cull_level_dim(ch, l + 1);
        x0 >>= 1; x1 >>= 1; y0 >>= 1; y1 >>= 1;
    }
    return m;
}
