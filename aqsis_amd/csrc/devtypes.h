// Plain-old-data records shared by the host launcher and the HIP kernels.
#pragma once

#include <stdint.h>

#ifndef RH_PRIM_SPHERE
#define RH_PRIM_SPHERE 1
#define RH_PRIM_BICUBIC 2
#define RH_PRIM_BILINEAR 3
#endif

// rh_prim.matte and rh_options.depth_filter values (include/aqsis_hider.h)
#define RH_DEV_MATTE 1
#define RH_DEV_MATTE_ALPHA 2
#define RH_DEV_DEPTH_MIDPOINT 1
#define RH_DEV_DEPTH_MAX 2
#define RH_DEV_DEPTH_AVERAGE 3
// StoreSample's modes (bucketprocessor.cpp:1076-1081, 1490-1535): the usual
// occluding hit; the midpoint store (a z display with the midpoint filter);
// every hit listed (a z display with the max or average filter)
#define RH_STORE_OCCLUDE 0
#define RH_STORE_MIDPOINT 1
#define RH_STORE_LIST_ALL 2

// GridDev::type flag of a deforming surface's grid (two motion keys): its
// parameters hold key 0's then key 1's, then the key times.
#define RH_GRID_MOTION 0x100

// Largest grid: u*v <= 256 micropolygons gives at most 2 x 257 vertices.
#define RH_MAX_GRID_VERTS 514
#define RH_MAX_GRID_ROWS 257
// Micropolygon record, in floats (the sampler rebuilds edge equations and
// the inverse-bilinear set-up from it when it stages a record):
//  [0..7]   raster x, y of the four vertices in grid order
//           (iu,iv), (iu+1,iv), (iu,iv+1), (iu+1,iv+1)
//  [8..11]  vertex depths z0..z3 (camera z, grid order)
//  [12..15] bound min x, max x, min y, max y
//  [16]     bound min z
//  [17]     flags (int bits): vertex-order code (bits 0..7), degenerate (8),
//           linear, i.e. one Newton step (9), opaque: Oi >= (1,1,1)
//           component-wise, color.h:432-435 (10)
//  [18]     grid index (int bits)  [19] unused
#define RH_MPREC_FLOATS 20
#define RH_MPREC_GRID 18
// Motion-frame micropolygon record, in floats:
//  [0..11]  key 0: x, y, z of the four vertices in grid order
//  [12..23] key 1 (a static micropolygon repeats key 0)
//  [24]     flags as above, plus RH_MB_MOVING  [25] grid index (int bits)
//  [26..27] key times  [28..31] static bound min x, max x, min y, max y
#define RH_MPREC_MB_FLOATS 32
#define RH_MPREC_MB_GRID 25
#define RH_MB_MOVING (1u << 11)
// Device counters: 0..4 the hider's SPL counters (sample tests, bound hits,
// stored hits, point-in-polygon hits, second Newton steps), 5 tasks whose
// list overflowed, 6 micropolygon-bucket references, 7 transparent-hit
// arena entries taken (may exceed the capacity: then the frame is redone
// with a larger arena), 8 unused.
#define RH_NCOUNTERS 9
// Hit-list entries (its occluding hit included) the combine kernel sorts in
// LDS; a sample with more transparent hits goes to the long-list pass.
#define RH_MAX_HITS 64
// Transparent-hit arena control words: 0 entries taken, 1 samples left to
// the long-list pass, 2 combine-list entries, 3 their hit-list entries,
// 4 and 5 the long-list pass's work offsets.
#define RH_HITCTL_WORDS 8
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
    float zThreshold[3];          // Option "limits" "zthreshold"
    uint32_t hitCapacity;         // transparent-hit arena entries
    float shutterOpen, shutterClose;
    float clipNear, clipFar;
    // depth of field (CqRenderer's lens state): on when dofOn != 0
    int dofOn;
    float dofMult, dofOneOverF;   // m_DofMultiplier, m_OneOverFocalDistance
    float dofScaleX, dofScaleY;   // m_DepthOfFieldScale
    // StoreSample's modes (bucketprocessor.cpp:1076-1081, 1499-1525):
    // cullableHits = 0 with a z display and the max or average depth filter
    // (every hit listed, no occlusion tests); midpointStore with a z display
    // and the midpoint filter (the occluding depth is the second nearest)
    int cullableHits, midpointStore;
    int depthFilter;              // RH_DEPTH_* for the combine
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
    int ntx, nty;                 // ... a row-major grid of ntx x nty
};

struct TaskDev {
    int bucket;
    int x0, y0, x1, y1;           // pixels, inside the bucket's owned region
    int pad0, pad1, pad2;
};

struct PrimDev {
    float Ci[3];
    float Oi[3];
    uint32_t constantShading;             // ShadingInterpolation "constant"
    uint32_t matte;                       // RH_MATTE_*
};

// Arguments of one sampling launch (see rh_sample in kernels.hip).
struct SampleLaunch {
    const FrameDev* fr;
    const BucketDev* buckets;
    const TaskDev* tasks;
    const uint2* items;                   // (task, first block | block count << 16)
    int itemBase, nStatic;                // static grids: items [itemBase, itemBase + nStatic)
    const uint32_t* nDevice;              // persistent: item count on the device
    int persistent;
    uint32_t* queueHead;
    const unsigned long long* capStart;
    const uint32_t* listCount;
    const uint32_t* lists;
    const float* mprec;
    const GridDev* grids;
    const PrimDev* prims;
    const float* pos2d;
    const float* pos1d;
    const uchar4* pixPattern;
    float4* samples;
    float4* hitData;                      // per transparent hit: u, v, depth, id
    uint32_t* hitNext;                    // per transparent hit: list link, surface
    uint32_t* hitCounter;
    uint4* combineList;
    float* frontDepth;
    uint32_t* rowCounts;
    int rowBase;
    uint32_t idCut;                       // only list entries below this id (0xffffffff: all)
    int depthOnly;                        // opaque depths only: no transparent hits are kept
    // Probing several buckets in one launch: per item (the static sampler's
    // items, the others' workgroups from itemBase) its own cut, replacing
    // idCut when set; for the one-workgroup-per-task samplers (motion, depth
    // of field) the task of each such workgroup (null: consecutive tasks).
    const uint32_t* itemCut;
    const uint32_t* taskList;
    // StoreSample's mode (RH_STORE_*, from FrameDev::cullableHits and
    // midpointStore): the static sampler's form for it
    int store;
};

// Lens tables of a frame with depth of field: per pattern the
// getShuffledIndices table (lens cell -> sample index) and the lens offsets
// projectToCircle(2p - 1); per extended pixel its lens-offset pattern; per
// lens cell the bound of its offsets (CalculateDofBounds: min x, min y,
// max x, max y).
struct DofTables {
    const uint8_t* shuffled;
    const float2* lens;
    const uint8_t* pixDof;
    const float4* cells;
};

// Sampling kernel variants: LDS chunk of CH records, SLOTS blocks per wave.
#define RH_SAMPLE_VARIANTS 4
#define RH_VARIANT_CH(v) ((v) < 2 ? 216 : 96)
#define RH_VARIANT_SLOTS(v) (((v) & 1) ? 2 : 4)
