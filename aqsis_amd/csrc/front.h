// Host front end of the hider: frame shape, sample plan and the split/dice
// schedule that the HIP kernels consume.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/aqsis_hider.h"
#include "hmath.h"

namespace rh {

// Frame shape: everything derived from the options alone.
struct Shape {
    rh_options o;
    int cropX0, cropX1, cropY0, cropY1;   // crop window in pixels
    int shiftX, shiftY;                   // DiscreteShift = lfloor(filterwidth/2)
    float xfwo2, yfwo2;                   // ceil(filterwidth)/2
    int nx, ny, ns;                       // PixelSamples and their product
    int nbx, nby;                         // buckets across the whole image
    int bx0, by0, bx1, by1;               // bucket region rendered (crop)
    int extX0, extY0, extW, extH;         // sampled extended region
    int width() const { return cropX1 - cropX0; }
    int height() const { return cropY1 - cropY0; }
    m44 camToRaster;                      // matSpaceToSpace("camera","raster")
    m44 diceM;                            // camToRaster with z column zeroed
    std::vector<float> filterTab;         // ((2sy+1)*(2sx+1)) * ns weights
    // Buckets of the region in row-major order; bucket k = (bx0 + k % nbr, by0 + k / nbr).
    int nbr, nbrY;                        // region width / height in buckets
    struct BucketInfo {
        int x, y, w, h;                   // display rectangle
        int sx0, sy0, sx1, sy1;           // owned sample region
    };
    std::vector<BucketInfo> buckets;
    int bucketIndex(int bx, int by) const { return (by - by0) * nbr + (bx - bx0); }
    // Device work split: a wave block is blockW x blockH pixels (64 samples
    // or one pixel); each bucket's owned region is cut into tasks of about
    // 4x4 blocks, one sampling workgroup each.  Tasks are in bucket order.
    int blockW = 1, blockH = 1;
    struct TaskInfo {
        int bucket;
        int x0, y0, x1, y1;               // pixels, inside the bucket's owned region
    };
    std::vector<TaskInfo> tasks;
    std::vector<int> taskStart;           // buckets + 1
};

bool make_shape(const rh_options& o, Shape& s, std::string& err);

// Sample plan: the pattern tables and per-pixel pattern choices that
// aqsis's global random stream produces for this frame shape.
struct Plan {
    int cache = 0;                        // patterns in the tables (250 or 1)
    std::vector<float> pos2d;             // cache * ns * 2
    std::vector<float> pos1d;             // cache * ns
    std::vector<uint8_t> pixPattern;      // ext pixels * 4: position, time, lod, shuffle
    std::vector<float> dither;            // display pixels: display 0's RandomFloat()
};

void make_plan(const Shape& s, Plan& p);

// GridRec::type flag of a deforming surface's grid: its parameters hold key
// 0's then key 1's (same layout), then the two key times.
constexpr int32_t GRID_MOTION = 0x100;

// One diced grid in dicing order.
struct GridRec {
    int32_t type, cu, cv, prim;
    uint32_t mpBase, vtxBase, paramOff;
    int32_t diceBucket;                   // region bucket index processed when diced
    int16_t cbx0, cby0, cbx1, cby1;       // candidate bucket range (inclusive)
    float px0, py0, px1, py1;             // raster bound with a 2-pixel guard band
};

struct Schedule {
    std::vector<GridRec> grids;
    std::vector<float> params;            // per-grid dicing parameters
    uint64_t micropolygons = 0;
    uint64_t vertices = 0;
    uint64_t surfaces = 0;
    bool motion = false;                  // some grid has two motion keys
    // Candidate grids per region bucket (CSR over the bucket list), in dicing order.
    std::vector<uint32_t> candStart;      // buckets + 1
    std::vector<uint32_t> candGrid;
    std::vector<uint32_t> segBucket;      // bucket of each candidate entry (a binning segment)
    std::vector<uint64_t> capStart;       // tasks + 1, per-task list capacity offsets
};

// Parameter layouts in Schedule::params (all floats):
//  sphere:   [0] radius, [1..16] object->camera matrix, [17] identity flag,
//            then sinT[cu+1], cosT[cu+1], sinP[cv+1], cosP[cv+1]
//  bicubic:  16 control points xyz (48 floats)
//  bilinear: 4 corners xyz (12 floats)
enum { SPH_RADIUS = 0, SPH_MAT = 1, SPH_IDENT = 17, SPH_TABLES = 18 };

bool make_schedule(const Shape& s, const rh_prim* prims, int n, Schedule& out, std::string& err,
                   int threads);

}  // namespace rh
