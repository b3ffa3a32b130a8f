// Plain-old-data records shared by the host launcher and the HIP kernels.
#pragma once

#include <stdint.h>

#ifndef RH_PRIM_SPHERE
#define RH_PRIM_SPHERE 1
#define RH_PRIM_BICUBIC 2
#define RH_PRIM_BILINEAR 3
#endif

// Largest grid: u*v <= 256 micropolygons gives at most 2 x 257 vertices.
#define RH_MAX_GRID_VERTS 514
#define RH_MAX_GRID_ROWS 257
// Hit-test record of one micropolygon, in floats:
//  [0..3] edge y-multipliers  [4..7] edge x-multipliers
//  [8..11] edge x origins     [12..15] edge y origins
//  [16..19] vertex depths z0..z3 (grid order)
//  [20..21] A  [22..23] E  [24..25] F  [26..27] G   (inverse bilinear)
//  [28..31] bound min x, max x, min y, max y
//  [32] bound min z  [33] linear flag (int bits)  [34] grid index (int bits)  [35] unused
#define RH_MPREC_FLOATS 36
// Device counters: 0..4 the hider's SPL counters (sample tests, bound hits,
// stored hits, point-in-polygon hits, second Newton steps), 5 tasks whose
// list overflowed, 6 micropolygon-bucket references.
#define RH_NCOUNTERS 8
// Sphere parameter layout (see front.h)
#define SPH_RADIUS_D 0
#define SPH_MAT_D 1
#define SPH_IDENT_D 17
#define SPH_TABLES_D 18

struct FrameDev {
    float camToRaster[16];
    int cropX0, cropX1, cropY0, cropY1;
    int shiftX, shiftY;
    float xfwo2, yfwo2;
    float filterX, filterY;       // PixelFilter widths
    int bucketX, bucketY;
    int bx0, by0, bx1, by1;       // bucket region
    int extX0, extY0, extW, extH;
    int ns, nx, ny;
    int blockW, blockH;           // pixels per 64-sample wave block
    float exposureGain, exposureGamma;
    float qOne, qMin, qMax, qDither;
};

struct GridDev {
    int type, cu, cv, prim;
    uint32_t mpBase, vtxBase, paramOff;
    int diceBucket;
};

struct BucketDev {
    int x, y, w, h;
    int sx0, sy0, sx1, sy1;       // owned sample region
    int bx, by;                   // bucket coordinates
    int task0, ntask;             // the bucket's sampling tasks
};

struct TaskDev {
    int bucket;
    int x0, y0, x1, y1;           // pixels, inside the bucket's owned region
    int pad0, pad1, pad2;
};

struct PrimDev {
    float Ci[3];
    float Oi[3];
    float pad[2];
};
