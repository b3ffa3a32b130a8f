// Host front end of the hider: frame shape, sample plan and the split/dice
// schedule that the HIP kernels consume.
#pragma once

#include <cmath>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/aqsis_hider.h"
#include "cullgeom.h"
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
    // Depth of field (RiDepthOfField; CqRenderer's lens state,
    // renderer.h:368-396): on when fstop < FLT_MAX.
    bool dof = false;
    float dofMult = 0, dofOneOverF = 0;   // m_DofMultiplier, m_OneOverFocalDistance
    float dofScaleX = 0, dofScaleY = 0;   // m_DepthOfFieldScale
    std::vector<float> dofCells;          // ns x (min x, min y, max x, max y): CalculateDofBounds
};

bool make_shape(const rh_options& o, Shape& s, std::string& err);

// GetCircleOfConfusion (renderer.h:401-406): raster radii at a camera depth.
inline void circle_of_confusion(const Shape& s, float depth, float& cx, float& cy) {
    const float c = s.dofMult * std::fabs(1.0f / depth - s.dofOneOverF);
    cx = s.dofScaleX * c;
    cy = s.dofScaleY * c;
}
// CqImagePixel::projectToCircle (imagepixel.h:429-436)
void project_to_circle(float x, float y, float& ox, float& oy);

// Sample plan: the pattern tables and per-pixel pattern choices that
// aqsis's global random stream produces for this frame shape.
struct Plan {
    int cache = 0;                        // patterns in the tables (250 or 1)
    std::vector<float> pos2d;             // cache * ns * 2
    std::vector<float> pos1d;             // cache * ns
    std::vector<uint8_t> pixPattern;      // ext pixels * 4: position, time, lod, shuffle
    std::vector<float> dither;            // display pixels: display 0's RandomFloat()
    // every display request's draws: request d's at [d * display pixels, ...)
    std::vector<float> ditherAll;
    // depth of field
    std::vector<uint8_t> shuffled;        // cache * ns: getShuffledIndices tables
    std::vector<float> lens;              // cache * ns * 2: projectToCircle(2 * pos2d - 1)
    std::vector<uint8_t> pixDof;          // ext pixels: the pattern of the lens offsets
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

// An occlusion test the final depths cannot settle (a surface's minimum
// depth within the margin of the leaves' deepest final depth, or a heap
// whose order does not bound the depths still to come): the tree's depths at
// the moment of the test decide it.
struct CullProbe {
    uint64_t pair;                        // (surface, bucket) pair of the test
    int32_t bucket;                       // region bucket index
    uint32_t mpCut;                       // micropolygons busted before the test
    float z;                              // the surface's minimum depth
    int32_t kind;                         // cull_cells: 0 no leaf, 1 cells, 2 cells + padding
    int32_t cx0, cx1, cy0, cy1;           // the test's sampled cells (inclusive)
    uint8_t culled;                       // the decision the schedule took
};

struct Schedule {
    std::vector<GridRec> grids;           // GridRec::paramOff: into the forest's parameters
    uint64_t paramFloats = 0;             // the forest's parameter floats (forest_params)
    uint64_t micropolygons = 0;
    uint64_t vertices = 0;
    uint64_t surfaces = 0;
    bool motion = false;                  // some grid has two motion keys
    // Candidate grids per region bucket (CSR over the bucket list), in dicing order.
    std::vector<uint32_t> candStart;      // buckets + 1
    std::vector<uint32_t> candGrid;
    std::vector<uint32_t> segBucket;      // bucket of each candidate entry (a binning segment)
    std::vector<uint64_t> capStart;       // tasks + 1, per-task list capacity offsets
    // occlusion culling (RenderSurface, bucketprocessor.cpp:945-957)
    uint64_t culled = 0;                  // surfaces culled by a bucket's occlusion tree
    std::vector<CullProbe> probes;        // tests left to the depths at their moment, in order
};

// The split trees of a scene's surfaces (CqSurface::Split / Diceable per
// primitive), built once and scheduled with or without occlusion culling.
struct Forest;
struct ForestFree { void operator()(Forest* f) const; };
using ForestPtr = std::unique_ptr<Forest, ForestFree>;

bool build_forest(const Shape& s, const rh_prim* prims, int n, ForestPtr& out, std::string& err, int threads);

// Every leaf's dicing parameters, in forest order: what a schedule's grids
// point into (GridRec::paramOff), the same for every schedule of the forest.
const float* forest_params(const Forest& f, uint64_t& n);

// The (surface, bucket) pairs an occlusion test may ask about: every surface
// with a cached raster bound, in every bucket of its bucket range (posting
// and reposting only ever move a surface inside it).
struct CullPairs {
    std::vector<float> bound;             // per surface: raster bound min x, min y, max x, max y
    std::vector<int32_t> range;           // per surface: region bucket range x0, y0, x1, y1 (empty: x1 < x0)
    std::vector<uint64_t> start;          // surfaces + 1: each surface's first pair
};
void cull_pairs(const Shape& s, const Forest& f, CullPairs& out);

// Each bucket's occlusion tree (setupTree) and depth-pyramid offsets; returns
// the pyramid floats of all buckets.
uint64_t cull_trees(const Shape& s, std::vector<CullTreeDev>& out);

// The answers a culling schedule takes: per pair, the deepest final depth of
// the leaves the surface's bound meets (negative: none), and decisions
// already fixed by probing (pair, culled), sorted by pair.
struct CullInput {
    const float* pairMax = nullptr;
    std::vector<std::pair<uint64_t, uint8_t>> fixed;
};

// The bucket loop of a forest: pop nearest first, cull (when `cull` is given
// and the surface's cull flag is set), dice or split, post the children.
bool simulate(const Shape& s, const rh_prim* prims, int n, const Forest& f, const CullInput* cull,
              const CullPairs* pairs, Schedule& out, std::string& err);

// The culling-free schedule's grids for an opaque-depth pass, without its
// bucket loop: every leaf diced in the first bucket of its bucket range, as
// the culling-free loop dices it (a surface is posted to its range's first
// bucket, which is never processed before it: roots are posted first, and a
// child's range lies inside its parent's, so its first bucket comes no
// earlier than the one its parent is split in), in forest order within a
// bucket.  Only the order within a bucket, and so the micropolygon ids,
// differ from simulate's.
bool depth_schedule(const Shape& s, const rh_prim* prims, int n, const Forest& f, Schedule& out, std::string& err);

// Relative margins around a cull test's depth inside which the final depths
// do not settle it.  A cull needs the final depths below the surface's
// minimum depth (and below the least one still to come from the heap) by
// more than micropolygon depths can fall below their surface's bound through
// dicing and bilinear rounding; keeping a surface needs them above it by more
// than the final depths of two dicing orders can differ (a bilinear depth
// below its micropolygon's least vertex depth, a few units in the last place).
constexpr float kCullMargin = 1.0f / 16384.0f;
constexpr float kKeepMargin = 1.0f / 65536.0f;

// Parameter layouts of a leaf in forest_params (all floats):
//  sphere:   [0] radius, [1..16] object->camera matrix, [17] identity flag,
//            then sinT[cu+1], cosT[cu+1], sinP[cv+1], cosP[cv+1]
//  bicubic:  16 control points xyz (48 floats)
//  bilinear: 4 corners xyz (12 floats)
enum { SPH_RADIUS = 0, SPH_MAT = 1, SPH_IDENT = 17, SPH_TABLES = 18 };

// StoreSample's isCullable without CSG (bucketprocessor.cpp:1076-1081):
// false with a z display and the max or average depth filter.
inline bool cullable_hits(const rh_options& o) {
    return !(o.display_z && (o.depth_filter == RH_DEPTH_MAX || o.depth_filter == RH_DEPTH_AVERAGE));
}
// Whether a surface's hits go to the per-sample lists (not the occluding
// hit): when hits are not cullable, or the surface is not opaque and not
// matte alpha (bucketprocessor.cpp:1490-1491).
inline bool hits_listed(const rh_options& o, const rh_prim& q) {
    if (!cullable_hits(o)) return true;
    return !(q.Os[0] >= 1.0f && q.Os[1] >= 1.0f && q.Os[2] >= 1.0f) && q.matte != RH_MATTE_ALPHA;
}

// build_forest + simulate without occlusion culling.
bool make_schedule(const Shape& s, const rh_prim* prims, int n, Schedule& out, std::string& err,
                   int threads);

}  // namespace rh
