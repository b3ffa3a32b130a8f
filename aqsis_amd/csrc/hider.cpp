// C ABI of the hider (include/aqsis_hider.h): context, device memory, the
// per-frame launch sequence and read-back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdint>
#include <mutex>
#include <cstring>
#include <list>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/aqsis_hider.h"
#include "devtypes.h"
#include "display.h"
#include "front.h"

extern "C" hipError_t rh_launch_dice(hipStream_t, const FrameDev*, const GridDev*, int, const PrimDev*,
                                     const float*, float*, float4*, float4*);
extern "C" hipError_t rh_launch_sample_mb(hipStream_t, const SampleLaunch*, int, const float4*);
extern "C" hipError_t rh_launch_bin(hipStream_t, const FrameDev*, const BucketDev*, int, const TaskDev*,
                                    const GridDev*, const uint32_t*, const uint32_t*, const uint32_t*, size_t,
                                    const unsigned long long*, const float4*, uint32_t*, uint32_t*, uint32_t*,
                                    uint32_t*, unsigned long long*, int, int, const float4*);
extern "C" hipError_t rh_launch_sample(hipStream_t, int, int, const SampleLaunch*);
extern "C" hipError_t rh_launch_sample_dof(hipStream_t, const SampleLaunch*, int, const DofTables*);
extern "C" hipError_t rh_sample_occupancy(int, int*);
extern "C" hipError_t rh_launch_plan_items(hipStream_t, const FrameDev*, const TaskDev*, int, const uint32_t*,
                                           uint32_t, int, uint2*, uint32_t*, uint32_t*);
extern "C" hipError_t rh_launch_frame_begin(hipStream_t, unsigned long long*, uint32_t*);
extern "C" hipError_t rh_launch_combine_long(hipStream_t, const FrameDev*, size_t, const uint4*, uint32_t*,
                                             const float4*, const uint32_t*, const PrimDev*, const float*, int, int,
                                             const GridDev*, int, float4*, float*, float*, uint32_t*, int*);
extern "C" hipError_t rh_launch_combine(hipStream_t, const FrameDev*, size_t, const uint4*, uint32_t*,
                                        const float4*, const uint32_t*, const PrimDev*, const float*, int, int,
                                        const GridDev*, int, float4*, float*);
extern "C" hipError_t rh_launch_reduce_counters(hipStream_t, const uint32_t*, int, const uint32_t*, int,
                                                unsigned long long*);
extern "C" size_t rh_filter_tile_lds(int, int, int);
extern "C" hipError_t rh_launch_box_masks(hipStream_t, const FrameDev*, size_t, const uchar4*, const float*,
                                          uint16_t*);
extern "C" hipError_t rh_launch_filter(hipStream_t, const FrameDev*, int, int, int, int, int, const float4*,
                                       const uint16_t*, const uchar4*, const float*, const float*, const float*,
                                       float*, uchar4*);
extern "C" hipError_t rh_launch_cull_pyramid(hipStream_t, const FrameDev*, const CullTreeDev*, int, const float4*, float*);
extern "C" hipError_t rh_launch_cull_pairs(hipStream_t, const CullTreeDev*, int, int, int, const float4*, const int4*,
                                           const unsigned long long*, int, unsigned long long, const float*, float*);
extern "C" hipError_t rh_launch_cull_probes_max(hipStream_t, const FrameDev*, const CullTreeDev*, const int4*,
                                                const int4*, int, const float4*, float*);

static_assert(rh::GRID_MOTION == RH_GRID_MOTION, "motion grid flag differs between host and device");
static_assert(RH_MATTE == RH_DEV_MATTE && RH_MATTE_ALPHA == RH_DEV_MATTE_ALPHA, "matte values differ");
static_assert(RH_DEPTH_MIDPOINT == RH_DEV_DEPTH_MIDPOINT && RH_DEPTH_MAX == RH_DEV_DEPTH_MAX &&
                  RH_DEPTH_AVERAGE == RH_DEV_DEPTH_AVERAGE, "depth filter values differ");

namespace {

thread_local std::string g_err;

rh_status fail(rh_status code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(RH_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));   \
    } while (0)

// A device buffer of a context, allocated in its stream's order from the
// device's memory pool: growing or freeing one does not wait for other
// contexts' work, and freed memory is reused without returning it to the
// driver (keep_pool_memory).
// Hand the device pool's unused memory back to the driver until the pool
// holds at most `keep` bytes.
void release_pool_memory(uint64_t keep) {
    int dev = 0;
    hipMemPool_t pool;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess)
        (void)hipMemPoolTrimTo(pool, (size_t)keep);
}

struct DBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipStream_t st = nullptr;             // the owning context's stream
    hipError_t ensure(size_t bytes) {
        if (bytes == 0) bytes = 16;
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipFreeAsync(p, st);
        p = nullptr;
        cap = 0;
        size_t n = bytes + bytes / 8 + 256;
        hipError_t e = hipMallocAsync(&p, n, st);
        if (e == hipErrorOutOfMemory) {
            // memory the pool holds unused (freed by other buffers) goes
            // back to the driver first
            (void)hipGetLastError();
            (void)hipStreamSynchronize(st);
            release_pool_memory(0);
            e = hipMallocAsync(&p, n, st);
        }
        if (e == hipSuccess) cap = n;
        return e;
    }
    void release() {
        if (p) (void)hipFreeAsync(p, st);
        p = nullptr;
        cap = 0;
    }
    template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

template <typename T> hipError_t upload(DBuf& b, const std::vector<T>& v, hipStream_t st) {
    hipError_t e = b.ensure(v.size() * sizeof(T));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st);
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

namespace rh {
// Error text of the host-only entry points compiled in other files.
void set_last_error(const std::string& msg) { g_err = msg; }
}  // namespace rh

struct rh_ctx {
    rh::Shape shape;
    rh::Plan plan;
    std::vector<rh_prim> prims;
    rh::Schedule sched;
    bool scheduled = false;
    int binTaskSide = 1;                  // largest task grid side of a bucket
    bool rendered = false;
    hipStream_t st = nullptr;
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    // frame shape
    DBuf dFrame, dBuckets, dTasks, dPos2d, dPos1d, dPix, dDither, dFilter, dBoxMask;
    DBuf dShuffled, dLens, dPixDof, dDofCells;   // depth of field: lens tables (DofTables)
    bool tiledFilter = false;
    // schedule
    DBuf dGrids, dParams, dPrims, dCandStart, dCandGrid, dCapStart, dSegBucket, dSegCounts;
    // per frame
    DBuf dMprec, dMpbound, dLists, dListCount, dSamples, dChannels, dRgba, dCounters, dBucketRefs;
    DBuf dMprecMB;                        // motion-frame micropolygon records
    // transparent hits: arena of (u, v, depth, id) and (link, surface) pairs
    DBuf dHitData, dHitNext, dHitCtl, dCombine, dFrontDepth;
    DBuf dLongWork;                       // the long-list combine's work arrays (grown on demand)
    bool transparent = false;             // some posted surface has Os < 1
    bool anyMatte = false;                // some posted surface is a matte (RH_MATTE)
    // work items: one per task (grid and per-bucket forms), cost-split (batched)
    DBuf dItemsStatic, dItems, dItemCount, dQueue, dRows;
    // occlusion culling: bucket trees, depth pyramids, (surface, bucket) tests, probe results
    DBuf dCullTrees, dCullPyr, dCullBounds, dCullRanges, dCullStart, dCullPairMax, dProbeOut;
    DBuf dProbeItems, dProbeTasks, dProbeCut, dProbeB, dProbeC;
    int maxItems = 0, persistentGrid = 0, maxRows = 0;
    rh_tuning tuning;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    int scheduleThreads = 0;              // host threads of the split/dice schedule (0: up to 16)
    void dropGraph() {
        if (graphExec) (void)hipGraphExecDestroy(graphExec);
        if (graph) (void)hipGraphDestroy(graph);
        graphExec = nullptr;
        graph = nullptr;
    }
    void* scratch = nullptr;              // this device's scratch array (rh_render_device_noalloc)
    FrameDev frame;                       // host copy of dFrame
    rh_stats stats;
    DofTables dof() const {
        DofTables t;
        t.shuffled = dShuffled.as<uint8_t>();
        t.lens = dLens.as<float2>();
        t.pixDof = dPixDof.as<uint8_t>();
        t.cells = dDofCells.as<float4>();
        return t;
    }
    template <typename F> void each_buffer(F f) {
        DBuf* all[] = {&dFrame, &dBuckets, &dTasks, &dPos2d, &dPos1d, &dPix, &dDither, &dFilter, &dBoxMask,
                       &dShuffled, &dLens, &dPixDof, &dDofCells,
                       &dGrids, &dParams,
                       &dPrims, &dCandStart, &dCandGrid, &dCapStart, &dMprec, &dMpbound, &dLists,
                       &dListCount, &dSamples, &dChannels, &dRgba, &dCounters, &dBucketRefs,
                       &dHitData, &dHitNext, &dHitCtl, &dCombine, &dFrontDepth, &dLongWork, &dItemsStatic, &dItems,
                       &dItemCount, &dQueue, &dRows, &dSegBucket, &dSegCounts, &dMprecMB, &dCullTrees, &dCullPyr,
                       &dCullBounds, &dCullRanges, &dCullStart, &dCullPairMax, &dProbeOut, &dProbeItems,
                       &dProbeTasks, &dProbeCut, &dProbeB, &dProbeC};
        for (DBuf* b : all) f(*b);
    }
    ~rh_ctx() {
        dropGraph();
        each_buffer([](DBuf& b) { b.release(); });
        if (st) (void)hipStreamSynchronize(st);   // the frees are ordered and done
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }
};

extern "C" {

void rh_options_default(rh_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->xres = 640; o->yres = 480;
    o->pixel_aspect = 1.0f;
    o->crop[0] = 0.0f; o->crop[1] = 1.0f; o->crop[2] = 0.0f; o->crop[3] = 1.0f;
    o->fov = 90.0f;
    o->clip_near = FLT_EPSILON; o->clip_far = FLT_MAX;
    o->xsamples = 2; o->ysamples = 2;
    o->filter = RH_FILTER_GAUSSIAN;
    o->filter_xwidth = 2.0f; o->filter_ywidth = 2.0f;
    o->exposure_gain = 1.0f; o->exposure_gamma = 1.0f;
    o->quantize_one = 255.0f; o->quantize_min = 0.0f; o->quantize_max = 255.0f; o->quantize_dither = 0.5f;
    o->shutter_open = 0.0f; o->shutter_close = 0.0f;
    o->bucket_x = 16; o->bucket_y = 16;
    o->eyesplits = 10;
    o->jitter = 1;
    o->num_displays = 1;
    o->zthreshold[0] = o->zthreshold[1] = o->zthreshold[2] = 1.0f;
    o->dof[0] = o->dof[1] = o->dof[2] = FLT_MAX;   // pinhole camera (options.cpp:297)
    o->depth_filter = RH_DEPTH_MIN;                 // optioncache.cpp:44
    o->display_z = 0;
}

const char* rh_last_error(void) { return g_err.c_str(); }

}  // extern "C"

// The device's memory pool keeps freed memory for reuse instead of handing
// it back to the driver at each synchronisation: contexts of different
// frame shapes come and go in the serving loop.
static void keep_pool_memory() {
    static std::once_flag once;
    std::call_once(once, [] {
        int dev = 0;
        hipMemPool_t pool;
        if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
            uint64_t keep = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
    });
}

// Sample plans of recent frame shapes.  A plan (the multijitter tables, the
// per-pixel pattern draws in bucket order and the dither draws) is a function
// of the frame shape alone, and drawing it is the dearest part of a context
// (about 0.24 s at 2048x2048, one sequential MT19937 stream); a renderer or server
// that re-creates contexts of shapes it has seen takes a copy instead.
// Least recently used plans go first beyond kPlanCacheBytes.
namespace {
using PlanKey = std::array<int, 14>;
struct PlanCache {
    std::mutex m;
    std::list<std::pair<PlanKey, std::shared_ptr<const rh::Plan>>> lru;   // front: most recent
    size_t bytes = 0;
};
PlanCache& plan_cache() {
    static PlanCache pc;
    return pc;
}
constexpr size_t kPlanCacheBytes = size_t(2) << 30;

size_t plan_bytes(const rh::Plan& p) {
    return (p.pos2d.size() + p.pos1d.size() + p.dither.size() + p.ditherAll.size() + p.lens.size()) * sizeof(float) +
           p.pixPattern.size() + p.pixDof.size() + p.shuffled.size();
}

// Everything make_plan reads from a shape.
PlanKey plan_key(const rh::Shape& s) {
    return {s.o.xres, s.o.yres, s.cropX0, s.cropX1, s.cropY0, s.cropY1, s.o.bucket_x, s.o.bucket_y,
            s.shiftX, s.shiftY, s.nx, s.ny, s.o.jitter ? 1 : 0, s.o.num_displays};
}

void cached_plan(const rh::Shape& s, rh::Plan& out) {
    const PlanKey key = plan_key(s);
    PlanCache& pc = plan_cache();
    {
        std::lock_guard<std::mutex> lk(pc.m);
        for (auto it = pc.lru.begin(); it != pc.lru.end(); ++it)
            if (it->first == key) {
                pc.lru.splice(pc.lru.begin(), pc.lru, it);
                out = *pc.lru.front().second;
                return;
            }
    }
    auto p = std::make_shared<rh::Plan>();
    rh::make_plan(s, *p);
    out = *p;
    const size_t b = plan_bytes(*p);
    if (b > kPlanCacheBytes / 4) return;
    std::lock_guard<std::mutex> lk(pc.m);
    for (const auto& e : pc.lru)
        if (e.first == key) return;       // another thread drew the same plan
    pc.lru.emplace_front(key, std::move(p));
    pc.bytes += b;
    while (pc.bytes > kPlanCacheBytes && !pc.lru.empty()) {
        pc.bytes -= plan_bytes(*pc.lru.back().second);
        pc.lru.pop_back();
    }
}
}  // namespace

// Device memory of rh_render_device_noalloc, for callers that cannot
// allocate: the frame's counters at 0, its hit-arena control words at 128,
// the sampling workgroups' counter rows from kScratchRowsOffset.  Every GPU
// holds its own copy until the process ends; host code reaches it only
// through the address hipGetSymbolAddress gives on that device.
static __device__ __align__(256) unsigned char scratch[RH_NOALLOC_SCRATCH_BYTES];
static constexpr size_t kScratchRowsOffset = 256;

namespace {
std::mutex g_scratchMu;
void* g_scratchAddr[256] = {};

// The scratch array's address on the current device, looked up once per device.
hipError_t scratch_address(void** out) {
    *out = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 256) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (!g_scratchAddr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratchAddr[dev] = p;
    }
    *out = g_scratchAddr[dev];
    return hipSuccess;
}
}  // namespace

extern "C" {

rh_status rh_create(const rh_options* opt, rh_ctx** out) {
    if (!opt || !out) return fail(RH_ERR_INVALID, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(RH_ERR_DEVICE, "no HIP device available");
    rh_ctx* c = new (std::nothrow) rh_ctx();
    if (!c) return fail(RH_ERR_NOMEM, "out of host memory");
    std::memset(&c->stats, 0, sizeof(c->stats));
    std::string err;
    if (!rh::make_shape(*opt, c->shape, err)) { delete c; return fail(RH_ERR_INVALID, err); }
    cached_plan(c->shape, c->plan);
    const rh::Shape& s = c->shape;
    keep_pool_memory();
    hipError_t e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
    c->each_buffer([c](DBuf& b) { b.st = c->st; });
    for (int i = 0; i < 6 && e == hipSuccess; ++i) e = hipEventCreate(&c->ev[i]);
    if (e != hipSuccess) { delete c; return fail(RH_ERR_DEVICE, std::string("stream/event creation: ") + hipGetErrorString(e)); }

    FrameDev f;
    std::memset(&f, 0, sizeof(f));
    for (int r = 0; r < 4; ++r)
        for (int cc = 0; cc < 4; ++cc) f.camToRaster[4 * r + cc] = s.camToRaster.e[r][cc];
    f.cropX0 = s.cropX0; f.cropX1 = s.cropX1; f.cropY0 = s.cropY0; f.cropY1 = s.cropY1;
    f.shiftX = s.shiftX; f.shiftY = s.shiftY; f.xfwo2 = s.xfwo2; f.yfwo2 = s.yfwo2;
    f.filterX = opt->filter_xwidth; f.filterY = opt->filter_ywidth;
    f.bucketX = opt->bucket_x; f.bucketY = opt->bucket_y;
    f.bx0 = s.bx0; f.by0 = s.by0; f.bx1 = s.bx1; f.by1 = s.by1;
    f.extX0 = s.extX0; f.extY0 = s.extY0; f.extW = s.extW; f.extH = s.extH;
    f.ns = s.ns; f.nx = s.nx; f.ny = s.ny;
    f.blockW = s.blockW; f.blockH = s.blockH;
    f.exposureGain = opt->exposure_gain; f.exposureGamma = opt->exposure_gamma;
    f.qOne = opt->quantize_one; f.qMin = opt->quantize_min; f.qMax = opt->quantize_max; f.qDither = opt->quantize_dither;
    for (int k = 0; k < 3; ++k) f.zThreshold[k] = opt->zthreshold[k];
    f.hitCapacity = 0;
    f.shutterOpen = opt->shutter_open; f.shutterClose = opt->shutter_close;
    f.clipNear = opt->clip_near; f.clipFar = opt->clip_far;
    f.dofOn = s.dof ? 1 : 0;
    f.dofMult = s.dofMult; f.dofOneOverF = s.dofOneOverF;
    f.dofScaleX = s.dofScaleX; f.dofScaleY = s.dofScaleY;
    f.cullableHits = rh::cullable_hits(*opt) ? 1 : 0;
    f.midpointStore = (opt->display_z && opt->depth_filter == RH_DEPTH_MIDPOINT) ? 1 : 0;
    f.depthFilter = opt->depth_filter;
    c->frame = f;

    std::vector<BucketDev> bk;
    c->binTaskSide = 1;
    for (size_t k = 0; k < s.buckets.size(); ++k) {
        const auto& b = s.buckets[k];
        BucketDev d;
        d.x = b.x; d.y = b.y; d.w = b.w; d.h = b.h;
        d.sx0 = b.sx0; d.sy0 = b.sy0; d.sx1 = b.sx1; d.sy1 = b.sy1;
        d.bx = s.bx0 + (int)(k % s.nbr); d.by = s.by0 + (int)(k / s.nbr);
        d.task0 = s.taskStart[k];
        d.ntask = s.taskStart[k + 1] - s.taskStart[k];
        // tasks are in row-major order: the first row's length is the width
        d.ntx = 0;
        while (d.ntx < d.ntask && s.tasks[d.task0 + d.ntx].y0 == s.tasks[d.task0].y0) ++d.ntx;
        d.nty = d.ntx > 0 ? d.ntask / d.ntx : 0;
        c->binTaskSide = std::max(c->binTaskSide, std::max(d.ntx, d.nty));
        bk.push_back(d);
    }
    std::vector<TaskDev> tk(s.tasks.size());
    for (size_t t = 0; t < s.tasks.size(); ++t) {
        const auto& T = s.tasks[t];
        tk[t].bucket = T.bucket;
        tk[t].x0 = T.x0; tk[t].y0 = T.y0; tk[t].x1 = T.x1; tk[t].y1 = T.y1;
        tk[t].pad0 = tk[t].pad1 = tk[t].pad2 = 0;
    }
    std::vector<FrameDev> fv(1, f);
    std::vector<unsigned long long> zero(RH_NCOUNTERS, 0);
    if ((e = upload(c->dFrame, fv, c->st)) != hipSuccess || (e = upload(c->dBuckets, bk, c->st)) != hipSuccess ||
        (e = upload(c->dTasks, tk, c->st)) != hipSuccess || (e = upload(c->dPos2d, c->plan.pos2d, c->st)) != hipSuccess ||
        (e = upload(c->dPos1d, c->plan.pos1d, c->st)) != hipSuccess || (e = upload(c->dPix, c->plan.pixPattern, c->st)) != hipSuccess ||
        (e = upload(c->dDither, c->plan.dither, c->st)) != hipSuccess || (e = upload(c->dFilter, s.filterTab, c->st)) != hipSuccess ||
        (e = c->dSamples.ensure((size_t)s.extW * s.extH * s.ns * 32)) != hipSuccess ||
        (e = c->dChannels.ensure((size_t)s.width() * s.height() * 9 * sizeof(float))) != hipSuccess ||
        (e = c->dRgba.ensure((size_t)s.width() * s.height() * 4)) != hipSuccess ||
        (e = c->dListCount.ensure(s.tasks.size() * sizeof(uint32_t))) != hipSuccess ||
        (e = c->dBucketRefs.ensure(s.buckets.size() * sizeof(uint32_t))) != hipSuccess ||
        (e = c->dHitCtl.ensure(RH_HITCTL_WORDS * sizeof(uint32_t))) != hipSuccess ||
        (e = upload(c->dCounters, zero, c->st)) != hipSuccess ||
        (s.dof && ((e = upload(c->dShuffled, c->plan.shuffled, c->st)) != hipSuccess ||
                   (e = upload(c->dLens, c->plan.lens, c->st)) != hipSuccess ||
                   (e = upload(c->dPixDof, c->plan.pixDof, c->st)) != hipSuccess ||
                   (e = upload(c->dDofCells, s.dofCells, c->st)) != hipSuccess))) {
        delete c;
        return fail(e == hipErrorOutOfMemory ? RH_ERR_NOMEM : RH_ERR_DEVICE, std::string("device set-up: ") + hipGetErrorString(e));
    }
    // The tiled filter reads each sample's box-test mask, fixed by the frame
    // shape's sample positions.
    c->tiledFilter = rh_filter_tile_lds(s.ns, s.shiftX, s.shiftY) != 0;
    if (c->tiledFilter) {
        const size_t nsamp = (size_t)s.extW * s.extH * s.ns;
        if ((e = c->dBoxMask.ensure(nsamp * sizeof(uint16_t))) != hipSuccess ||
            (e = rh_launch_box_masks(c->st, c->dFrame.as<FrameDev>(), nsamp, c->dPix.as<uchar4>(),
                                     c->dPos2d.as<float>(), c->dBoxMask.as<uint16_t>())) != hipSuccess) {
            delete c;
            return fail(e == hipErrorOutOfMemory ? RH_ERR_NOMEM : RH_ERR_DEVICE,
                        std::string("device set-up: ") + hipGetErrorString(e));
        }
    }
    // Work items: one per task for the static forms; up to one per block
    // for the batched form.
    {
        std::vector<uint2> it(s.tasks.size());
        int total = 0;
        for (size_t t = 0; t < s.tasks.size(); ++t) {
            const auto& T = s.tasks[t];
            const int nblk = ((T.x1 - T.x0 + s.blockW - 1) / s.blockW) * ((T.y1 - T.y0 + s.blockH - 1) / s.blockH);
            it[t] = make_uint2((uint32_t)t, (uint32_t)nblk << 16);
            total += nblk;
        }
        c->maxItems = std::max(total, 1);
        int cus = 0, dev = 0, occ = 0;
        if ((e = hipGetDevice(&dev)) == hipSuccess)
            e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        int best = 1;
        for (int v = 0; v < RH_SAMPLE_VARIANTS && e == hipSuccess; ++v) {
            e = rh_sample_occupancy(v, &occ);
            best = std::max(best, occ);
        }
        c->persistentGrid = std::max(1, cus * best);
        const int unit = 8 * 16;
        const int staticGrid = ((int)s.tasks.size() + unit - 1) / unit * unit;
        c->maxRows = std::max(staticGrid, c->persistentGrid);
        if (e == hipSuccess) e = upload(c->dItemsStatic, it, c->st);
        if (e == hipSuccess) e = c->dItems.ensure((size_t)c->maxItems * sizeof(uint2));
        if (e == hipSuccess) e = c->dItemCount.ensure(16);
        if (e == hipSuccess) e = c->dQueue.ensure(16);
        if (e == hipSuccess) e = c->dRows.ensure((size_t)c->maxRows * 5 * sizeof(uint32_t));
    }
    // The scratch array's address on this device (for the no-allocation
    // form only: a failed lookup leaves it null, and that form reports it).
    if (e == hipSuccess && scratch_address(&c->scratch) != hipSuccess) {
        c->scratch = nullptr;
        (void)hipGetLastError();
    }
    c->tuning.mode = RH_MODE_GRID;
    c->tuning.lds_records = 216;
    c->tuning.blocks_per_wave = 4;
    c->tuning.split_cost = 1u << 20;
    c->tuning.frame_ms = 0.0;
    if (e != hipSuccess || (e = hipStreamSynchronize(c->st)) != hipSuccess) {
        delete c;
        return fail(RH_ERR_DEVICE, std::string("device set-up: ") + hipGetErrorString(e));
    }
    *out = c;
    return RH_OK;
}

rh_status rh_post_surfaces(rh_ctx* c, const rh_prim* prims, int32_t n) {
    if (!c || (n > 0 && !prims) || n < 0) return fail(RH_ERR_INVALID, "bad arguments");
    c->prims.insert(c->prims.end(), prims, prims + n);
    c->scheduled = false;
    c->rendered = false;
    return RH_OK;
}

rh_status rh_clear_surfaces(rh_ctx* c) {
    if (!c) return fail(RH_ERR_INVALID, "null context");
    c->prims.clear();
    c->scheduled = false;
    c->rendered = false;
    return RH_OK;
}

// Transparent-hit arena entries a frame is likely to need: each transparent
// micropolygon covers about one pixel at ShadingRate 1, so about ns samples
// (at least 1M entries).  A frame that needs more is redone once with the
// exact size.
static uint64_t arena_guess(const rh::Shape& s, const rh::Schedule& S, const rh_prim* prims) {
    uint64_t tmp = 0;
    for (const auto& g : S.grids)
        if (rh::hits_listed(s.o, prims[g.prim])) tmp += (uint64_t)g.cu * g.cv;
    if (tmp == 0) return 0;
    return std::min<uint64_t>(std::max<uint64_t>(1u << 20, tmp * (uint64_t)s.ns), 0x7fffffffull);
}

// Size the transparent-hit arena and publish its capacity to the kernels.
static rh_status set_hit_capacity(rh_ctx* c, uint64_t n) {
    if (n < 64) n = 64;
    if (n > 0x7fffffffull) return fail(RH_ERR_CAPACITY, "transparent hits exceed 2^31 entries");
    HIP_TRY(c->dHitData.ensure(n * sizeof(float4)));
    HIP_TRY(c->dHitNext.ensure(n * 2 * sizeof(uint32_t)));
    c->frame.hitCapacity = (uint32_t)n;
    c->dropGraph();
    std::vector<FrameDev> fv(1, c->frame);
    HIP_TRY(upload(c->dFrame, fv, c->st));
    return RH_OK;
}

static rh_status upload_schedule(rh_ctx* c);
static rh_status cull_schedule(rh_ctx* c, const rh::Forest& f);

// Some surface may be occlusion culled (Attribute "cull" "hidden", default 1;
// RenderSurface tests none when hits are not cullable, bucketprocessor.cpp:945-947).
static bool any_cullable(const rh_options& o, const std::vector<rh_prim>& prims) {
    if (!rh::cullable_hits(o)) return false;
    for (const auto& q : prims)
        if (!q.no_cull_hidden) return true;
    return false;
}

// The frame's split/dice schedule: the culling-free one first, then, when
// surfaces may be culled, the schedule the reference's occlusion culling
// gives (cull_schedule).
static rh_status install_schedule(rh_ctx* c, rh::Schedule&& S, const rh::Forest& f,
                                  std::chrono::steady_clock::time_point t0) {
    c->sched = std::move(S);
    c->stats.cull_passes = 1;
    c->stats.cull_probes = 0;
    c->stats.ms_cull = 0.0;
    c->stats.ms_cull_depth = 0.0;
    c->stats.ms_cull_probes = 0.0;
    c->stats.ms_schedule = ms_since(t0);
    // every schedule of the forest dices from its parameter table
    uint64_t nparams = 0;
    const float* params = rh::forest_params(f, nparams);
    HIP_TRY(c->dParams.ensure(nparams * sizeof(float)));
    if (nparams) HIP_TRY(hipMemcpyAsync(c->dParams.p, params, nparams * sizeof(float), hipMemcpyHostToDevice, c->st));
    rh_status r = upload_schedule(c);
    if (r != RH_OK || c->prims.empty() || !any_cullable(c->shape.o, c->prims)) return r;
    const auto tc = std::chrono::steady_clock::now();
    r = cull_schedule(c, f);
    c->stats.ms_cull = ms_since(tc);
    c->stats.ms_schedule = ms_since(t0);
    return r;
}

static rh_status schedule(rh_ctx* c) {
    c->dropGraph();
    auto t0 = std::chrono::steady_clock::now();
    std::string err;
    int threads = (int)std::thread::hardware_concurrency();
    if (threads > 16) threads = 16;
    if (c->scheduleThreads > 0) threads = c->scheduleThreads;
    rh::ForestPtr f;
    rh::Schedule S;
    if (!rh::build_forest(c->shape, c->prims.data(), (int)c->prims.size(), f, err, threads))
        return fail(RH_ERR_INVALID, err);
    c->stats.ms_forest = ms_since(t0);
    const auto ts = std::chrono::steady_clock::now();
    // A scene that may be occlusion culled first renders the culling-free
    // grids for their opaque depths, in any order within a bucket.
    if (!(any_cullable(c->shape.o, c->prims)
              ? rh::depth_schedule(c->shape, c->prims.data(), (int)c->prims.size(), *f, S, err)
              : rh::simulate(c->shape, c->prims.data(), (int)c->prims.size(), *f, nullptr, nullptr, S, err)))
        return fail(RH_ERR_INVALID, err);
    c->stats.ms_simulate = ms_since(ts);
    return install_schedule(c, std::move(S), *f, t0);
}

// Device copies of a computed schedule.
static rh_status upload_schedule(rh_ctx* c) {
    c->dropGraph();
    auto t1 = std::chrono::steady_clock::now();
    const rh::Schedule& S = c->sched;
    std::vector<GridDev> g(S.grids.size());
    for (size_t i = 0; i < g.size(); ++i) {
        const auto& r = S.grids[i];
        g[i].type = r.type; g[i].cu = r.cu; g[i].cv = r.cv; g[i].prim = r.prim;
        g[i].mpBase = r.mpBase; g[i].vtxBase = r.vtxBase; g[i].paramOff = r.paramOff; g[i].diceBucket = r.diceBucket;
    }
    std::vector<PrimDev> p(c->prims.size());
    for (size_t i = 0; i < p.size(); ++i) {
        for (int k = 0; k < 3; ++k) {
            // constant shader: Oi = Os; Ci = Os * Cs (shaders/surface/constant.sl)
            p[i].Oi[k] = c->prims[i].Os[k];
            p[i].Ci[k] = c->prims[i].Os[k] * c->prims[i].Cs[k];
        }
        p[i].constantShading = c->prims[i].shading_interpolation == RH_SHADING_CONSTANT ? 1u : 0u;
        p[i].matte = (uint32_t)c->prims[i].matte;
    }
    std::vector<unsigned long long> cap(S.capStart.begin(), S.capStart.end());
    HIP_TRY(upload(c->dGrids, g, c->st));
    HIP_TRY(upload(c->dPrims, p, c->st));
    HIP_TRY(upload(c->dCandStart, S.candStart, c->st));
    HIP_TRY(upload(c->dCandGrid, S.candGrid, c->st));
    HIP_TRY(upload(c->dSegBucket, S.segBucket, c->st));
    HIP_TRY(c->dSegCounts.ensure(S.candGrid.size() * 17 * sizeof(uint32_t)));
    HIP_TRY(upload(c->dCapStart, cap, c->st));
    HIP_TRY(c->dMprec.ensure(S.micropolygons * RH_MPREC_FLOATS * sizeof(float)));
    HIP_TRY(c->dMpbound.ensure(S.micropolygons * sizeof(float4)));
    HIP_TRY(c->dLists.ensure((size_t)S.capStart.back() * sizeof(uint32_t)));
    if (S.motion) HIP_TRY(c->dMprecMB.ensure(S.micropolygons * RH_MPREC_MB_FLOATS * sizeof(float)));
    // Transparent hits need list storage; start at two per sample and grow
    // when a frame asks for more.
    const rh::Shape& s = c->shape;
    bool transparent = false, anyMatte = false;
    for (const auto& q : c->prims) {
        if (rh::hits_listed(s.o, q)) transparent = true;
        if (q.matte == RH_MATTE) anyMatte = true;
    }
    c->transparent = transparent;
    c->anyMatte = anyMatte;
    const uint64_t nsamp = (uint64_t)s.extW * s.extH * s.ns;
    const uint64_t want = transparent ? arena_guess(s, S, c->prims.data()) : 64;
    HIP_TRY(c->dCombine.ensure((transparent ? nsamp : 1) * sizeof(uint4)));
    HIP_TRY(c->dFrontDepth.ensure((transparent ? nsamp : 1) * sizeof(float)));
    if (want > c->frame.hitCapacity) {
        rh_status r = set_hit_capacity(c, want);
        if (r != RH_OK) return r;
    }
    HIP_TRY(hipStreamSynchronize(c->st));
    c->stats.ms_upload = ms_since(t1);
    c->scheduled = true;
    return RH_OK;
}

static int variant_of(const rh_tuning& t) {
    return (t.lds_records == 216 ? 0 : 2) + (t.blocks_per_wave == 4 ? 0 : 1);
}

// Where a frame keeps its counters (RH_NCOUNTERS), its hit-arena control
// words (4) and the per-workgroup counter rows (5 per row).
struct FrameCtl {
    unsigned long long* counters;
    uint32_t* hitCtl;
    uint32_t* rows;
};

static FrameCtl context_ctl(rh_ctx* c) {
    return FrameCtl{c->dCounters.as<unsigned long long>(), c->dHitCtl.as<uint32_t>(), c->dRows.as<uint32_t>()};
}

// The filter + expose + quantise stage of a frame.
static hipError_t enqueue_filter(rh_ctx* c) {
    const rh::Shape& s = c->shape;
    return rh_launch_filter(c->st, c->dFrame.as<FrameDev>(), s.width(), s.height(), s.ns, s.shiftX, s.shiftY,
                            c->dSamples.as<float4>(), c->tiledFilter ? c->dBoxMask.as<uint16_t>() : nullptr,
                            c->dPix.as<uchar4>(), c->dPos2d.as<float>(), c->dFilter.as<float>(),
                            c->dDither.as<float>(), c->dChannels.as<float>(), c->dRgba.as<uchar4>());
}

// Whether the combine takes its full form: a depth filter other than min, or
// some matte surface.
static int combine_full(const rh_ctx* c) {
    return (c->frame.depthFilter != RH_DEPTH_MIN || c->anyMatte) ? 1 : 0;
}

// StoreSample's mode of a frame (RH_STORE_*).
static int store_mode(const FrameDev& f) {
    if (!f.cullableHits) return RH_STORE_LIST_ALL;
    return f.midpointStore ? RH_STORE_MIDPOINT : RH_STORE_OCCLUDE;
}

// The micropolygon records the combine finds an occluding hit's grid in:
// the motion records in a frame with moving micropolygons; null when no
// surface is a matte (the combine then looks up no surface for it).
static const float* occl_records(rh_ctx* c) {
    if (!c->anyMatte) return nullptr;
    return c->sched.motion ? reinterpret_cast<const float*>(c->dMprecMB.p) : c->dMprec.as<float>();
}

// Enqueue one device frame on the context's stream.  With `timed`, HIP
// events bracket the stages (not inside a graph capture).  `gridOnly`
// takes the static grid layout whatever the tuning says.
static hipError_t enqueue_frame(rh_ctx* c, bool timed, const FrameCtl& k, bool gridOnly = false) {
    const rh::Schedule& S = c->sched;
    const rh::Shape& s = c->shape;
    const FrameDev* f = c->dFrame.as<FrameDev>();
    const int nb = (int)s.buckets.size(), nt = (int)s.tasks.size();
    rh_tuning tu = c->tuning;
    if (gridOnly) tu.mode = RH_MODE_GRID;
    hipError_t e;
#define ENQ(x) do { if ((e = (x)) != hipSuccess) return e; } while (0)
    if (timed) ENQ(hipEventRecord(c->ev[0], c->st));
    ENQ(rh_launch_frame_begin(c->st, k.counters, k.hitCtl));
    ENQ(rh_launch_dice(c->st, f, c->dGrids.as<GridDev>(), (int)S.grids.size(), c->dPrims.as<PrimDev>(),
                       c->dParams.as<float>(), c->dMprec.as<float>(), c->dMpbound.as<float4>(),
                       S.motion ? c->dMprecMB.as<float4>() : nullptr));
    if (timed) ENQ(hipEventRecord(c->ev[1], c->st));
    ENQ(rh_launch_bin(c->st, f, c->dBuckets.as<BucketDev>(), nb, c->dTasks.as<TaskDev>(), c->dGrids.as<GridDev>(),
                      c->dCandStart.as<uint32_t>(), c->dCandGrid.as<uint32_t>(), c->dSegBucket.as<uint32_t>(),
                      S.candGrid.size(), c->dCapStart.as<unsigned long long>(), c->dMpbound.as<float4>(),
                      c->dSegCounts.as<uint32_t>(), c->dLists.as<uint32_t>(), c->dListCount.as<uint32_t>(),
                      c->dBucketRefs.as<uint32_t>(), k.counters, c->binTaskSide,
                      ((c->frame.bucketX & (c->frame.bucketX - 1)) == 0 && (c->frame.bucketY & (c->frame.bucketY - 1)) == 0) ? 1 : 0,
                      S.motion ? c->dMprecMB.as<float4>() : nullptr));
    if (timed) ENQ(hipEventRecord(c->ev[2], c->st));
    SampleLaunch a;
    a.fr = f; a.buckets = c->dBuckets.as<BucketDev>(); a.tasks = c->dTasks.as<TaskDev>();
    a.items = c->dItemsStatic.as<uint2>(); a.itemBase = 0; a.nStatic = nt; a.nDevice = c->dItemCount.as<uint32_t>();
    a.persistent = 0; a.queueHead = c->dQueue.as<uint32_t>();
    a.capStart = c->dCapStart.as<unsigned long long>(); a.listCount = c->dListCount.as<uint32_t>();
    a.lists = c->dLists.as<uint32_t>(); a.mprec = c->dMprec.as<float>(); a.grids = c->dGrids.as<GridDev>();
    a.prims = c->dPrims.as<PrimDev>(); a.pos2d = c->dPos2d.as<float>(); a.pos1d = c->dPos1d.as<float>();
    a.pixPattern = c->dPix.as<uchar4>(); a.samples = c->dSamples.as<float4>(); a.hitData = c->dHitData.as<float4>();
    a.hitNext = c->dHitNext.as<uint32_t>(); a.hitCounter = k.hitCtl;
    a.combineList = c->dCombine.as<uint4>(); a.frontDepth = c->transparent ? c->dFrontDepth.as<float>() : nullptr;
    a.rowCounts = k.rows; a.rowBase = 0;
    a.idCut = 0xffffffffu; a.depthOnly = 0;
    a.itemCut = nullptr; a.taskList = nullptr;
    a.store = store_mode(c->frame);
    const int variant = variant_of(tu);
    int rows = 0;
    if (S.motion) {
        // Frames with moving micropolygons take the motion form of the
        // sampler, whatever the layout.
        ENQ(rh_launch_sample_mb(c->st, &a, nt, c->dMprecMB.as<float4>()));
        rows = nt;
    } else if (s.dof) {
        // Depth of field: the lens form of the sampler, whatever the layout.
        const DofTables dt = c->dof();
        ENQ(rh_launch_sample_dof(c->st, &a, nt, &dt));
        rows = nt;
    } else if (tu.mode == RH_MODE_PLAIN) {
        // One launch per bucket, in bucket order.
        for (int b = 0; b < nb; ++b) {
            a.itemBase = s.taskStart[b];
            a.nStatic = s.taskStart[b + 1] - s.taskStart[b];
            a.rowBase = s.taskStart[b];
            ENQ(rh_launch_sample(c->st, variant, a.nStatic, &a));
        }
        rows = nt;
    } else if (tu.mode == RH_MODE_BATCHED) {
        ENQ(rh_launch_plan_items(c->st, f, c->dTasks.as<TaskDev>(), nt, c->dListCount.as<uint32_t>(), tu.split_cost,
                                 4 * tu.blocks_per_wave, c->dItems.as<uint2>(), c->dItemCount.as<uint32_t>(),
                                 c->dQueue.as<uint32_t>()));
        a.items = c->dItems.as<uint2>();
        a.persistent = 1;
        ENQ(rh_launch_sample(c->st, variant, c->persistentGrid, &a));
        rows = c->persistentGrid;
    } else {
        const int unit = 8 * 16;
        rows = (nt + unit - 1) / unit * unit;
        ENQ(rh_launch_sample(c->st, variant, rows, &a));
    }
    if (timed) ENQ(hipEventRecord(c->ev[5], c->st));
    if (c->transparent)
        ENQ(rh_launch_combine(c->st, f, (size_t)s.extW * s.extH * s.ns, c->dCombine.as<uint4>(),
                              k.hitCtl, c->dHitData.as<float4>(), c->dHitNext.as<uint32_t>(),
                              c->dPrims.as<PrimDev>(), occl_records(c), S.motion ? RH_MPREC_MB_FLOATS : RH_MPREC_FLOATS,
                              S.motion ? RH_MPREC_MB_GRID : RH_MPREC_GRID, c->dGrids.as<GridDev>(), combine_full(c),
                              c->dSamples.as<float4>(),
                              c->dFrontDepth.as<float>()));
    if (timed) ENQ(hipEventRecord(c->ev[3], c->st));
    ENQ(rh_launch_reduce_counters(c->st, k.rows, rows, c->dBucketRefs.as<uint32_t>(), nb, k.counters));
    ENQ(enqueue_filter(c));
    if (timed) ENQ(hipEventRecord(c->ev[4], c->st));
#undef ENQ
    return hipSuccess;
}

// Samples with more transparent hits than the combine kernel sorts in LDS:
// the long-list combine over work arrays sized from the frame's counts
// (hc[1] samples, hc[3] list entries), then the filter again.
static rh_status combine_long_lists(rh_ctx* c, const FrameCtl& k, const uint32_t* hc) {
    const rh::Shape& s = c->shape;
    const size_t entries = hc[3], deferred = hc[1];
    const size_t bytes = entries * (sizeof(float) + sizeof(uint32_t)) + deferred * 192 * sizeof(int);
    HIP_TRY(c->dLongWork.ensure(bytes));
    unsigned char* w = static_cast<unsigned char*>(c->dLongWork.p);
    float* workD = reinterpret_cast<float*>(w);
    uint32_t* workX = reinterpret_cast<uint32_t*>(w + entries * sizeof(float));
    int* workStk = reinterpret_cast<int*>(w + entries * (sizeof(float) + sizeof(uint32_t)));
    const bool motion = c->sched.motion;
    HIP_TRY(rh_launch_combine_long(c->st, c->dFrame.as<FrameDev>(), (size_t)s.extW * s.extH * s.ns,
                                   c->dCombine.as<uint4>(), k.hitCtl, c->dHitData.as<float4>(),
                                   c->dHitNext.as<uint32_t>(), c->dPrims.as<PrimDev>(), occl_records(c),
                                   motion ? RH_MPREC_MB_FLOATS : RH_MPREC_FLOATS, motion ? RH_MPREC_MB_GRID : RH_MPREC_GRID,
                                   c->dGrids.as<GridDev>(), combine_full(c), c->dSamples.as<float4>(),
                                   c->dFrontDepth.as<float>(), workD, workX, workStk));
    HIP_TRY(enqueue_filter(c));
    HIP_TRY(hipStreamSynchronize(c->st));
    return RH_OK;
}

// The sampling launch arguments of a frame's depth pass: opaque depths only
// (no transparent hit is kept), written to the frame's sample records.
static SampleLaunch depth_launch(rh_ctx* c) {
    SampleLaunch a;
    std::memset(&a, 0, sizeof(a));
    a.fr = c->dFrame.as<FrameDev>(); a.buckets = c->dBuckets.as<BucketDev>(); a.tasks = c->dTasks.as<TaskDev>();
    a.items = c->dItemsStatic.as<uint2>(); a.itemBase = 0; a.nStatic = (int)c->shape.tasks.size();
    a.nDevice = c->dItemCount.as<uint32_t>(); a.persistent = 0; a.queueHead = c->dQueue.as<uint32_t>();
    a.capStart = c->dCapStart.as<unsigned long long>(); a.listCount = c->dListCount.as<uint32_t>();
    a.lists = c->dLists.as<uint32_t>(); a.mprec = c->dMprec.as<float>(); a.grids = c->dGrids.as<GridDev>();
    a.prims = c->dPrims.as<PrimDev>(); a.pos2d = c->dPos2d.as<float>(); a.pos1d = c->dPos1d.as<float>();
    a.pixPattern = c->dPix.as<uchar4>(); a.samples = c->dSamples.as<float4>(); a.hitData = c->dHitData.as<float4>();
    a.hitNext = c->dHitNext.as<uint32_t>(); a.hitCounter = c->dHitCtl.as<uint32_t>();
    a.combineList = c->dCombine.as<uint4>(); a.frontDepth = nullptr;
    a.rowCounts = c->dRows.as<uint32_t>(); a.rowBase = 0;
    a.idCut = 0xffffffffu; a.depthOnly = 1;
    // with the midpoint store the occluding depth is the second nearest
    a.store = store_mode(c->frame);
    return a;
}

// Dice, bust and bin the resident schedule (its task lists), then, with
// `sample`, sample every task for opaque depths.
static hipError_t enqueue_depth_pass(rh_ctx* c, bool sample) {
    const rh::Schedule& S = c->sched;
    const rh::Shape& s = c->shape;
    const FrameDev* f = c->dFrame.as<FrameDev>();
    hipError_t e;
    if ((e = rh_launch_frame_begin(c->st, c->dCounters.as<unsigned long long>(), c->dHitCtl.as<uint32_t>())) != hipSuccess)
        return e;
    if ((e = rh_launch_dice(c->st, f, c->dGrids.as<GridDev>(), (int)S.grids.size(), c->dPrims.as<PrimDev>(),
                            c->dParams.as<float>(), c->dMprec.as<float>(), c->dMpbound.as<float4>(),
                            S.motion ? c->dMprecMB.as<float4>() : nullptr)) != hipSuccess)
        return e;
    if ((e = rh_launch_bin(c->st, f, c->dBuckets.as<BucketDev>(), (int)s.buckets.size(), c->dTasks.as<TaskDev>(),
                           c->dGrids.as<GridDev>(), c->dCandStart.as<uint32_t>(), c->dCandGrid.as<uint32_t>(),
                           c->dSegBucket.as<uint32_t>(), S.candGrid.size(), c->dCapStart.as<unsigned long long>(),
                           c->dMpbound.as<float4>(), c->dSegCounts.as<uint32_t>(), c->dLists.as<uint32_t>(),
                           c->dListCount.as<uint32_t>(), c->dBucketRefs.as<uint32_t>(),
                           c->dCounters.as<unsigned long long>(), c->binTaskSide,
                           ((c->frame.bucketX & (c->frame.bucketX - 1)) == 0 &&
                            (c->frame.bucketY & (c->frame.bucketY - 1)) == 0) ? 1 : 0,
                           S.motion ? c->dMprecMB.as<float4>() : nullptr)) != hipSuccess)
        return e;
    if (!sample) return hipSuccess;
    SampleLaunch a = depth_launch(c);
    const int nt = (int)s.tasks.size();
    if (S.motion) return rh_launch_sample_mb(c->st, &a, nt, c->dMprecMB.as<float4>());
    if (s.dof) {
        const DofTables dt = c->dof();
        return rh_launch_sample_dof(c->st, &a, nt, &dt);
    }
    return rh_launch_sample(c->st, 0, nt, &a);
}

// The schedule aqsis's occlusion culling gives (RenderSurface,
// bucketprocessor.cpp:945-957; RepostSurface, imagebuffer.cpp:449-483;
// CqOcclusionTree::canCull, occlusion.cpp:143-196).  A bucket's tree holds
// the opaque depths of its samples at the moment a surface is taken from
// its heap; depths only decrease, and a surface popped later than another
// cannot reach below that one's minimum depth except by rounding, so the
// frame's final depths (from the culling-free schedule, which the resident
// buffers hold) decide every test whose depths lie clear of the surface's
// minimum depth by a margin.  The others (CullProbe) are decided by the
// bucket's depths at that moment: the bucket's lists are rendered up to the
// test's first micropolygon and the leaves read back.  A probe that
// overturns a decision fixes it and the schedule is computed again.
static rh_status cull_schedule(rh_ctx* c, const rh::Forest& forest) {
    const rh::Shape& s = c->shape;
    const int n = (int)c->prims.size();
    const auto td = std::chrono::steady_clock::now();
    // 1) the final opaque depths of the culling-free schedule
    HIP_TRY(enqueue_depth_pass(c, true));
    // 2) every bucket's depth pyramid and every (surface, bucket) test
    rh::CullPairs pairs;
    rh::cull_pairs(s, forest, pairs);
    std::vector<CullTreeDev> trees;
    const uint64_t pyrFloats = rh::cull_trees(s, trees);
    const size_t nnodes = pairs.start.size() - 1;
    const uint64_t npairs = pairs.start.back();
    std::vector<float> pairMax(npairs, -1.0f);
    if (npairs > 0) {
        HIP_TRY(upload(c->dCullTrees, trees, c->st));
        HIP_TRY(c->dCullPyr.ensure(pyrFloats * sizeof(float)));
        HIP_TRY(rh_launch_cull_pyramid(c->st, c->dFrame.as<FrameDev>(), c->dCullTrees.as<CullTreeDev>(),
                                       (int)trees.size(), c->dSamples.as<float4>(), c->dCullPyr.as<float>()));
        std::vector<unsigned long long> st(pairs.start.begin(), pairs.start.end());
        HIP_TRY(upload(c->dCullBounds, pairs.bound, c->st));
        HIP_TRY(upload(c->dCullRanges, pairs.range, c->st));
        HIP_TRY(upload(c->dCullStart, st, c->st));
        HIP_TRY(c->dCullPairMax.ensure(npairs * sizeof(float)));
        HIP_TRY(rh_launch_cull_pairs(c->st, c->dCullTrees.as<CullTreeDev>(), s.bx0, s.by0, s.nbr,
                                     c->dCullBounds.as<float4>(), c->dCullRanges.as<int4>(),
                                     c->dCullStart.as<unsigned long long>(), (int)nnodes, npairs,
                                     c->dCullPyr.as<float>(), c->dCullPairMax.as<float>()));
        HIP_TRY(hipMemcpyAsync(pairMax.data(), c->dCullPairMax.p, npairs * sizeof(float), hipMemcpyDeviceToHost, c->st));
        HIP_TRY(hipStreamSynchronize(c->st));
    }
    c->stats.ms_cull_depth = ms_since(td);
    // 3) the culling schedule, its probes, and again while a probe overturns a decision
    rh::CullInput in;
    in.pairMax = pairMax.data();
    std::string err;
    for (int pass = 0;; ++pass) {
        rh::Schedule S;
        const auto ts = std::chrono::steady_clock::now();
        if (!rh::simulate(s, c->prims.data(), n, forest, &in, &pairs, S, err)) return fail(RH_ERR_INVALID, err);
        c->stats.ms_simulate += ms_since(ts);
        c->sched = std::move(S);
        rh_status r = upload_schedule(c);
        if (r != RH_OK) return r;
        c->stats.cull_passes++;
        const std::vector<rh::CullProbe>& probes = c->sched.probes;
        if (probes.empty()) break;
        if (pass > 64) return fail(RH_ERR_STATE, "occlusion culling: probes kept overturning decisions");
        const auto tp = std::chrono::steady_clock::now();
        // the schedule's task lists, then each probe's bucket up to its moment.
        // Probes go in rounds, round r holding each bucket's r-th probe: a
        // round's buckets are sampled in one launch, each task cut at its
        // probe's moment, and one launch reads all the round's leaves.
        HIP_TRY(enqueue_depth_pass(c, false));
        const size_t np = probes.size();
        std::vector<int> nth(s.buckets.size(), 0), round(np);
        int nrounds = 0;
        for (size_t i = 0; i < np; ++i) {
            round[i] = nth[probes[i].bucket]++;
            nrounds = std::max(nrounds, round[i] + 1);
        }
        std::vector<std::vector<uint32_t>> byRound(nrounds);
        for (size_t i = 0; i < np; ++i) byRound[round[i]].push_back((uint32_t)i);
        std::vector<uint2> items;                 // static sampler: (task, all its blocks)
        std::vector<uint32_t> taskOf, cut;        // per item: task, list cut
        std::vector<int4> pb, pc;                 // per probe in round order: (bucket, kind, result), cells
        std::vector<int> itemStart(nrounds + 1, 0), probeStart(nrounds + 1, 0);
        for (int r = 0; r < nrounds; ++r) {
            for (uint32_t i : byRound[r]) {
                const rh::CullProbe& pr = probes[i];
                const int t0 = s.taskStart[pr.bucket], t1 = s.taskStart[pr.bucket + 1];
                for (int t = t0; t < t1; ++t) {
                    const auto& T = s.tasks[t];
                    const int nblk = ((T.x1 - T.x0 + s.blockW - 1) / s.blockW) * ((T.y1 - T.y0 + s.blockH - 1) / s.blockH);
                    items.push_back(make_uint2((uint32_t)t, (uint32_t)nblk << 16));
                    taskOf.push_back((uint32_t)t);
                    cut.push_back(pr.mpCut);
                }
                pb.push_back(make_int4(pr.bucket, t1 > t0 ? pr.kind : 0, (int)i, 0));
                pc.push_back(make_int4(pr.cx0, pr.cx1, pr.cy0, pr.cy1));
            }
            itemStart[r + 1] = (int)items.size();
            probeStart[r + 1] = (int)pb.size();
        }
        HIP_TRY(c->dProbeOut.ensure(np * sizeof(float)));
        HIP_TRY(upload(c->dProbeItems, items, c->st));
        HIP_TRY(upload(c->dProbeTasks, taskOf, c->st));
        HIP_TRY(upload(c->dProbeCut, cut, c->st));
        HIP_TRY(upload(c->dProbeB, pb, c->st));
        HIP_TRY(upload(c->dProbeC, pc, c->st));
        for (int r = 0; r < nrounds; ++r) {
            const int ni = itemStart[r + 1] - itemStart[r];
            if (ni > 0) {
                SampleLaunch a = depth_launch(c);
                a.items = c->dProbeItems.as<uint2>();
                a.itemBase = itemStart[r];
                a.nStatic = ni;
                a.rowBase = 0;
                a.itemCut = c->dProbeCut.as<uint32_t>();
                a.taskList = c->dProbeTasks.as<uint32_t>();
                if (c->sched.motion) {
                    HIP_TRY(rh_launch_sample_mb(c->st, &a, ni, c->dMprecMB.as<float4>()));
                } else if (s.dof) {
                    const DofTables dt = c->dof();
                    HIP_TRY(rh_launch_sample_dof(c->st, &a, ni, &dt));
                } else {
                    HIP_TRY(rh_launch_sample(c->st, 0, ni, &a));
                }
            }
            HIP_TRY(rh_launch_cull_probes_max(c->st, c->dFrame.as<FrameDev>(), c->dCullTrees.as<CullTreeDev>(),
                                              c->dProbeB.as<int4>() + probeStart[r],
                                              c->dProbeC.as<int4>() + probeStart[r],
                                              probeStart[r + 1] - probeStart[r], c->dSamples.as<float4>(),
                                              c->dProbeOut.as<float>()));
        }
        std::vector<float> res(probes.size());
        HIP_TRY(hipMemcpyAsync(res.data(), c->dProbeOut.p, res.size() * sizeof(float), hipMemcpyDeviceToHost, c->st));
        HIP_TRY(hipStreamSynchronize(c->st));
        c->stats.ms_cull_probes += ms_since(tp);
        c->stats.cull_probes += probes.size();
        // Decisions up to the first overturned one stand (the schedule before
        // it is the reference's); that one is fixed and the rest recomputed.
        bool overturned = false;
        for (size_t i = 0; i < probes.size() && !overturned; ++i) {
            const uint8_t exact = res[i] < probes[i].z ? 1 : 0;
            in.fixed.emplace_back(probes[i].pair, exact);
            overturned = exact != probes[i].culled;
        }
        std::sort(in.fixed.begin(), in.fixed.end());
        if (!overturned) break;
    }
    // the scratch of the tests goes back to the pool
    for (DBuf* b : {&c->dCullPyr, &c->dCullBounds, &c->dCullRanges, &c->dCullStart, &c->dCullPairMax,
                    &c->dProbeItems, &c->dProbeTasks, &c->dProbeCut, &c->dProbeB, &c->dProbeC})
        b->release();
    return RH_OK;
}

static rh_status finish_frame(rh_ctx* c, const unsigned long long* cnt, bool staged);

// One frame on the resident inputs.  The batched form replays its frame
// graph unless `eager` asks for the same launches issued one by one with a
// timestamp between the stages.
static rh_status render_frame(rh_ctx* c, bool eager) {
    if (!c) return fail(RH_ERR_INVALID, "null context");
    if (!c->scheduled) return fail(RH_ERR_STATE, "rh_render must run before rh_render_device");
    // The kernels overwrite the frame's buffers: until this frame completes,
    // nothing may be read back.
    c->rendered = false;
    const rh::Schedule& S = c->sched;
    const bool graphed = c->tuning.mode == RH_MODE_BATCHED && !eager;
    unsigned long long cnt[RH_NCOUNTERS];
    for (int attempt = 0;; ++attempt) {
        if (graphed) {
            // The batched form replays one captured graph per frame.
            if (!c->graphExec) {
                // Only the whole frame is timed in this form: events recorded
                // inside a captured graph do not report elapsed times here.
                HIP_TRY(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
                hipError_t e = enqueue_frame(c, false, context_ctl(c));
                hipGraph_t g = nullptr;
                hipError_t e2 = hipStreamEndCapture(c->st, &g);
                if (e != hipSuccess || e2 != hipSuccess) {
                    if (g) (void)hipGraphDestroy(g);
                    return fail(RH_ERR_DEVICE, std::string("frame graph capture: ") +
                                                   hipGetErrorString(e != hipSuccess ? e : e2));
                }
                c->graph = g;
                HIP_TRY(hipGraphInstantiate(&c->graphExec, c->graph, nullptr, nullptr, 0));
            }
            HIP_TRY(hipEventRecord(c->ev[0], c->st));
            HIP_TRY(hipGraphLaunch(c->graphExec, c->st));
            HIP_TRY(hipEventRecord(c->ev[4], c->st));
        } else {
            HIP_TRY(enqueue_frame(c, true, context_ctl(c)));
        }
        HIP_TRY(hipEventSynchronize(c->ev[4]));
        HIP_TRY(hipMemcpy(cnt, c->dCounters.p, sizeof(cnt), hipMemcpyDeviceToHost));
        uint32_t hc[4];
        HIP_TRY(hipMemcpy(hc, c->dHitCtl.p, sizeof(hc), hipMemcpyDeviceToHost));
        if (hc[0] > c->frame.hitCapacity) {
            // The arena ran out: redo the frame with room for every hit.
            if (attempt >= 2) return fail(RH_ERR_CAPACITY, "transparent-hit arena kept overflowing");
            rh_status r = set_hit_capacity(c, (uint64_t)hc[0] + hc[0] / 4 + 64);
            if (r != RH_OK) return r;
            continue;
        }
        rh_status r = finish_frame(c, cnt, !graphed);
        if (r != RH_OK || hc[1] == 0) return r;
        // samples with more hits than the combine kernel holds
        c->rendered = false;
        const auto t0 = std::chrono::steady_clock::now();
        r = combine_long_lists(c, context_ctl(c), hc);
        if (r != RH_OK) return r;
        const double ms = ms_since(t0);
        c->stats.ms_composite += ms;
        c->stats.ms_device += ms;
        c->rendered = true;
        return RH_OK;
    }
}

// A completed frame's counters and stage times into the context's stats.
static rh_status finish_frame(rh_ctx* c, const unsigned long long* cnt, bool staged) {
    const rh::Schedule& S = c->sched;
    float t[4] = {0.0f, 0.0f, 0.0f, 0.0f}, tComp = 0.0f;
    if (staged) {
        for (int i = 0; i < 4; ++i) HIP_TRY(hipEventElapsedTime(&t[i], c->ev[i], c->ev[i + 1]));
        HIP_TRY(hipEventElapsedTime(&tComp, c->ev[5], c->ev[3]));
    }
    c->stats.spl_count = cnt[0];
    c->stats.spl_bound_hits = cnt[1];
    c->stats.spl_hits = cnt[2];
    c->stats.inside = cnt[3];
    c->stats.inside_nonlinear = cnt[4];
    c->stats.grids = S.grids.size();
    c->stats.micropolygons = S.micropolygons;
    c->stats.surfaces = S.surfaces;
    c->stats.occlusion_culled = S.culled;
    c->stats.ms_dice = t[0];
    c->stats.ms_bin = t[1];
    c->stats.ms_sample = t[2];
    c->stats.ms_filter = t[3];
    c->stats.ms_composite = tComp;
    float total = 0;
    HIP_TRY(hipEventElapsedTime(&total, c->ev[0], c->ev[4]));
    c->stats.ms_device = total;
    c->stats.mp_bucket_refs = cnt[6];
    if (cnt[5] != 0)
        return fail(RH_ERR_CAPACITY, "internal: a task's micropolygon list exceeded its capacity (" +
                                         std::to_string(cnt[5]) + " tasks)");
    c->rendered = true;
    return RH_OK;
}

rh_status rh_render_device(rh_ctx* c) { return render_frame(c, false); }

// rh_render_device without device allocation: the frame's counters, hit
// control words and counter rows in the GPU's scratch array, every other
// buffer the context's (sized by rh_render), the static grid layout, issued
// eagerly.  A frame that would need a larger transparent-hit arena fails
// instead of growing it.
rh_status rh_render_device_noalloc(rh_ctx* c) {
    if (!c) return fail(RH_ERR_INVALID, "null context");
    if (!c->scheduled) return fail(RH_ERR_STATE, "rh_render must run before rh_render_device_noalloc");
    if (!c->scratch) return fail(RH_ERR_STATE, "no scratch memory on this context's device");
    const int nt = (int)c->shape.tasks.size();
    const size_t rows = (c->sched.motion || c->shape.dof) ? (size_t)nt : (size_t)((nt + 127) / 128 * 128);
    if (kScratchRowsOffset + rows * 5 * sizeof(uint32_t) > RH_NOALLOC_SCRATCH_BYTES)
        return fail(RH_ERR_CAPACITY, "the frame has " + std::to_string(nt) + " sampling tasks; the no-allocation "
                                     "form holds the counters of at most " +
                                     std::to_string((RH_NOALLOC_SCRATCH_BYTES - kScratchRowsOffset) / 20) + " workgroups");
    unsigned char* base = static_cast<unsigned char*>(c->scratch);
    const FrameCtl k{reinterpret_cast<unsigned long long*>(base), reinterpret_cast<uint32_t*>(base + 128),
                     reinterpret_cast<uint32_t*>(base + kScratchRowsOffset)};
    c->rendered = false;
    HIP_TRY(enqueue_frame(c, true, k, true));
    HIP_TRY(hipEventSynchronize(c->ev[4]));
    unsigned long long cnt[RH_NCOUNTERS];
    uint32_t hc[4];
    HIP_TRY(hipMemcpy(cnt, k.counters, sizeof(cnt), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hc, k.hitCtl, sizeof(hc), hipMemcpyDeviceToHost));
    if (hc[0] > c->frame.hitCapacity)
        return fail(RH_ERR_CAPACITY, "the frame needs " + std::to_string(hc[0]) + " transparent-hit entries and the "
                                     "context holds " + std::to_string(c->frame.hitCapacity) +
                                     " (rh_render_device grows the arena)");
    if (hc[1] != 0)
        return fail(RH_ERR_CAPACITY, std::to_string(hc[1]) + " samples collected more than " +
                                         std::to_string(RH_MAX_HITS - 1) + " transparent hits: their combine needs "
                                         "work memory (rh_render_device allocates it)");
    return finish_frame(c, cnt, true);
}

rh_status rh_render_stages(rh_ctx* c) { return render_frame(c, true); }

rh_status rh_set_tuning(rh_ctx* c, const rh_tuning* t) {
    if (!c || !t) return fail(RH_ERR_INVALID, "null argument");
    if (t->mode != RH_MODE_GRID && t->mode != RH_MODE_PLAIN && t->mode != RH_MODE_BATCHED)
        return fail(RH_ERR_INVALID, "unknown mode");
    if (t->lds_records != 216 && t->lds_records != 96) return fail(RH_ERR_INVALID, "lds_records must be 216 or 96");
    if (t->blocks_per_wave != 4 && t->blocks_per_wave != 2)
        return fail(RH_ERR_INVALID, "blocks_per_wave must be 4 or 2");
    if (t->split_cost == 0) return fail(RH_ERR_INVALID, "split_cost must be positive");
    const double ms = c->tuning.frame_ms;
    c->tuning = *t;
    c->tuning.frame_ms = ms;
    c->dropGraph();
    return RH_OK;
}

rh_status rh_get_tuning(const rh_ctx* c, rh_tuning* t) {
    if (!c || !t) return fail(RH_ERR_INVALID, "null argument");
    *t = c->tuning;
    return RH_OK;
}

rh_status rh_autotune(rh_ctx* c, rh_tuning* chosen) {
    if (!c) return fail(RH_ERR_INVALID, "null context");
    if (!c->scheduled) {
        rh_status r = schedule(c);
        if (r != RH_OK) return r;
    }
    rh_tuning best;
    double bestMs = 1e300;
    const int lds[2] = {216, 96}, bpw[2] = {4, 2};
    // 0: the static grid.  Small split costs make many short items whose list
    // re-reads outweigh the balance they buy on dense frames; large ones
    // leave a few heavy buckets as the tail.
    const uint32_t split[5] = {0, 1u << 16, 1u << 18, 1u << 20, 1u << 22};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 5; ++k) {
                rh_tuning t;
                t.mode = split[k] ? RH_MODE_BATCHED : RH_MODE_GRID;
                t.lds_records = lds[i];
                t.blocks_per_wave = bpw[j];
                t.split_cost = split[k] ? split[k] : (1u << 16);
                t.frame_ms = 0.0;
                rh_status r = rh_set_tuning(c, &t);
                // first frame captures the graph; the second is timed
                for (int rep = 0; rep < 2 && r == RH_OK; ++rep) r = rh_render_device(c);
                if (r != RH_OK) return r;
                if (c->stats.ms_device < bestMs) {
                    bestMs = c->stats.ms_device;
                    best = t;
                }
            }
    best.frame_ms = bestMs;
    rh_status r = rh_set_tuning(c, &best);
    if (r != RH_OK) return r;
    c->tuning.frame_ms = bestMs;
    if (chosen) *chosen = c->tuning;
    return RH_OK;
}

rh_status rh_render(rh_ctx* c) {
    if (!c) return fail(RH_ERR_INVALID, "null context");
    if (!c->scheduled) {
        rh_status r = schedule(c);
        if (r != RH_OK) return r;
    }
    return rh_render_device(c);
}

rh_status rh_frame_size(const rh_ctx* c, int32_t* w, int32_t* h, int32_t* ew, int32_t* eh, int32_t* spp) {
    if (!c) return fail(RH_ERR_INVALID, "null context");
    if (w) *w = c->shape.width();
    if (h) *h = c->shape.height();
    if (ew) *ew = c->shape.extW;
    if (eh) *eh = c->shape.extH;
    if (spp) *spp = c->shape.ns;
    return RH_OK;
}

rh_status rh_read_channels(rh_ctx* c, float* out) {
    if (!c || !out) return fail(RH_ERR_INVALID, "bad arguments");
    if (!c->rendered) return fail(RH_ERR_STATE, "no rendered frame");
    HIP_TRY(hipMemcpy(out, c->dChannels.p, (size_t)c->shape.width() * c->shape.height() * 9 * sizeof(float),
                      hipMemcpyDeviceToHost));
    return RH_OK;
}

rh_status rh_read_rgba8(rh_ctx* c, uint8_t* out) {
    if (!c || !out) return fail(RH_ERR_INVALID, "bad arguments");
    if (!c->rendered) return fail(RH_ERR_STATE, "no rendered frame");
    HIP_TRY(hipMemcpy(out, c->dRgba.p, (size_t)c->shape.width() * c->shape.height() * 4, hipMemcpyDeviceToHost));
    return RH_OK;
}

rh_status rh_display_format(const rh_options* opt, int32_t* type, int32_t* bytes) {
    if (!opt) return fail(RH_ERR_INVALID, "null options");
    const int t = rh::display_type(*opt);
    if (type) *type = t;
    if (bytes) *bytes = rh::display_type_bytes(t);
    return RH_OK;
}

rh_status rh_format_display(const rh_options* opt, const float* channels, const float* dither, int64_t npix,
                            int32_t type, void* out) {
    if (!opt || npix < 0 || (npix > 0 && (!channels || !dither || !out)))
        return fail(RH_ERR_INVALID, "bad arguments");
    const int vb = rh::display_type_bytes(type);
    if (vb == 0) return fail(RH_ERR_INVALID, "unknown display value type " + std::to_string(type));
    unsigned char* dst = static_cast<unsigned char*>(out);
    for (int64_t i = 0; i < npix; ++i)
        rh::format_pixel(*opt, channels + i * 9, (double)dither[i], type, dst + i * 4 * vb);
    return RH_OK;
}

// The frame's display values of `type`, width x height x 4 values.
static rh_status display_values(rh_ctx* c, int32_t type, std::vector<unsigned char>& out) {
    const rh::Shape& s = c->shape;
    const size_t npix = (size_t)s.width() * s.height();
    const int vb = rh::display_type_bytes(type);
    if (vb == 0) return fail(RH_ERR_INVALID, "unknown display value type " + std::to_string(type));
    out.resize(npix * 4 * vb);
    if (type == RH_DSPY_UNSIGNED8 && rh::display_type(s.o) == RH_DSPY_UNSIGNED8) {
        // the device quantised these already
        HIP_TRY(hipMemcpy(out.data(), c->dRgba.p, npix * 4, hipMemcpyDeviceToHost));
        return RH_OK;
    }
    std::vector<float> ch(npix * 9);
    HIP_TRY(hipMemcpy(ch.data(), c->dChannels.p, ch.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < npix; ++i)
        rh::format_pixel(s.o, &ch[i * 9], (double)c->plan.dither[i], type, &out[i * 4 * vb]);
    return RH_OK;
}

rh_status rh_select_format(float one, float min, float max, int32_t* type) {
    if (!type) return fail(RH_ERR_INVALID, "null type");
    *type = rh::select_data_format(one, min, max);
    return RH_OK;
}

rh_status rh_read_display_request(rh_ctx* c, int32_t mode, const float* quantize, int32_t index, int32_t type,
                                  void* out) {
    if (!c || !out || !quantize) return fail(RH_ERR_INVALID, "bad arguments");
    if (!c->rendered) return fail(RH_ERR_STATE, "no rendered frame");
    const rh::Shape& s = c->shape;
    const int nv = rh::mode_channels(mode);
    if (nv == 0) return fail(RH_ERR_INVALID, "mode must combine RH_MODE_RGB, RH_MODE_A and RH_MODE_Z");
    if (index < 0 || index >= std::max(1, s.o.num_displays))
        return fail(RH_ERR_INVALID, "display index outside the frame's display requests");
    const int vb = rh::display_type_bytes(type);
    if (vb == 0) return fail(RH_ERR_INVALID, "unknown display value type " + std::to_string(type));
    const size_t npix = (size_t)s.width() * s.height();
    std::vector<float> ch(npix * 9);
    HIP_TRY(hipMemcpy(ch.data(), c->dChannels.p, ch.size() * sizeof(float), hipMemcpyDeviceToHost));
    const float* d = &c->plan.ditherAll[(size_t)index * npix];
    unsigned char* dst = static_cast<unsigned char*>(out);
    for (size_t i = 0; i < npix; ++i) dst += rh::format_request(quantize, mode, &ch[i * 9], (double)d[i], type, dst);
    return RH_OK;
}

rh_status rh_read_display(rh_ctx* c, int32_t type, void* out) {
    if (!c || !out) return fail(RH_ERR_INVALID, "bad arguments");
    if (!c->rendered) return fail(RH_ERR_STATE, "no rendered frame");
    std::vector<unsigned char> v;
    rh_status r = display_values(c, type, v);
    if (r != RH_OK) return r;
    std::memcpy(out, v.data(), v.size());
    return RH_OK;
}

rh_status rh_display_as(rh_ctx* c, rh_dspy_data_fn data, void* image, int32_t scanline_order, int32_t type) {
    if (!c || !data) return fail(RH_ERR_INVALID, "bad arguments");
    if (!c->rendered) return fail(RH_ERR_STATE, "no rendered frame");
    const rh::Shape& s = c->shape;
    const int W = s.width();
    std::vector<unsigned char> img;
    rh_status r = display_values(c, type, img);
    if (r != RH_OK) return r;
    const size_t entry = 4 * (size_t)rh::display_type_bytes(type);
    // a pixel (x, y) of the crop window, in raster coordinates
    auto at = [&](int x, int y) { return &img[((size_t)(y - s.cropY0) * W + (x - s.cropX0)) * entry]; };
    std::vector<unsigned char> bucket, rows;
    if (scanline_order) rows.resize((size_t)W * entry * s.o.bucket_y);
    for (const auto& b : s.buckets) {
        // the bucket's display rectangle inside the crop window
        const int x0 = std::max(b.x, s.cropX0), x1 = std::min(b.x + b.w, s.cropX1);
        const int y0 = std::max(b.y, s.cropY0), y1 = std::min(b.y + b.h, s.cropY1);
        if (x1 <= x0 || y1 <= y0) continue;
        const size_t span = (size_t)(x1 - x0) * entry;
        if (!scanline_order) {
            // the rectangle's pixels, row by row (FormatBucketForDisplay)
            bucket.resize(span * (y1 - y0));
            for (int y = y0; y < y1; ++y) std::memcpy(&bucket[(size_t)(y - y0) * span], at(x0, y), span);
            if (data(image, x0, x1, y0, y1, (int)entry, bucket.data()) != 0)
                return fail(RH_ERR_STATE, "the display driver stopped the delivery");
            continue;
        }
        // CollapseBucketsToScanlines: gather the bucket row, send whole rows
        // when its last bucket has arrived (SendToDisplay)
        for (int y = y0; y < y1; ++y)
            std::memcpy(&rows[((size_t)(y - y0) * W + (x0 - s.cropX0)) * entry], at(x0, y), span);
        if (x1 >= s.cropX1)
            for (int y = y0; y < y1; ++y)
                if (data(image, s.cropX0, s.cropX1, y, y + 1, (int)entry, &rows[(size_t)(y - y0) * W * entry]) != 0)
                    return fail(RH_ERR_STATE, "the display driver stopped the delivery");
    }
    return RH_OK;
}

rh_status rh_display(rh_ctx* c, rh_dspy_data_fn data, void* image, int32_t scanline_order) {
    if (!c) return fail(RH_ERR_INVALID, "null context");
    return rh_display_as(c, data, image, scanline_order, rh::display_type(c->shape.o));
}

rh_status rh_read_samples(rh_ctx* c, uint32_t* mp_id, float* depth, float* rgbaz) {
    if (!c) return fail(RH_ERR_INVALID, "null context");
    if (!c->rendered) return fail(RH_ERR_STATE, "no rendered frame");
    const size_t n = (size_t)c->shape.extW * c->shape.extH * c->shape.ns;
    std::vector<float> buf(n * 8), front;
    HIP_TRY(hipMemcpy(buf.data(), c->dSamples.p, n * 32, hipMemcpyDeviceToHost));
    if (c->transparent && depth) {
        // depth of the front-most hit, kept apart for combined samples
        front.resize(n);
        HIP_TRY(hipMemcpy(front.data(), c->dFrontDepth.p, n * sizeof(float), hipMemcpyDeviceToHost));
    }
    for (size_t i = 0; i < n; ++i) {
        const float* r = &buf[i * 8];
        uint32_t id;
        std::memcpy(&id, &r[7], 4);
        if (mp_id) mp_id[i] = id;
        if (depth) depth[i] = front.empty() ? r[3] : front[i];
        if (rgbaz) {
            float* o = rgbaz + i * 7;
            o[0] = r[0]; o[1] = r[1]; o[2] = r[2]; o[3] = r[4]; o[4] = r[5]; o[5] = r[6]; o[6] = r[3];
        }
    }
    return RH_OK;
}

rh_status rh_get_stats(const rh_ctx* c, rh_stats* out) {
    if (!c || !out) return fail(RH_ERR_INVALID, "bad arguments");
    *out = c->stats;
    return RH_OK;
}

void rh_destroy(rh_ctx* c) { delete c; }

}  // extern "C"

// Used by the serving loop (server.cpp): hand the device pool's unused memory
// back to the driver until the pool holds at most `keep` bytes.  The pool
// keeps freed memory for reuse; a server trims it to its budget after each
// job and when it destroys contexts, so that neither its buffers nor what the
// driver allocates itself (kernel scratch at launch) find the device full of
// freed blocks.
void rh_internal_trim_pool(uint64_t keep) { release_pool_memory(keep); }

uint64_t rh_internal_device_bytes(const rh_ctx* c) {
    const DBuf* all[] = {&c->dFrame, &c->dBuckets, &c->dTasks, &c->dPos2d, &c->dPos1d, &c->dPix, &c->dDither,
                         &c->dFilter, &c->dBoxMask, &c->dGrids, &c->dParams, &c->dPrims, &c->dCandStart,
                         &c->dCandGrid, &c->dCapStart, &c->dSegBucket, &c->dSegCounts, &c->dMprec, &c->dMpbound, &c->dLists,
                         &c->dListCount, &c->dSamples, &c->dChannels, &c->dRgba, &c->dCounters,
                         &c->dBucketRefs, &c->dHitData, &c->dHitNext, &c->dHitCtl, &c->dCombine, &c->dFrontDepth,
                         &c->dItemsStatic, &c->dItems, &c->dItemCount, &c->dQueue, &c->dRows, &c->dMprecMB,
                         &c->dCullTrees, &c->dCullPyr, &c->dCullBounds, &c->dCullRanges, &c->dCullStart,
                         &c->dCullPairMax, &c->dProbeOut, &c->dProbeItems, &c->dProbeTasks, &c->dProbeCut,
                         &c->dProbeB, &c->dProbeC, &c->dShuffled, &c->dLens, &c->dPixDof, &c->dDofCells,
                         &c->dLongWork};
    uint64_t total = 0;
    for (const DBuf* b : all) total += b->cap;
    return total;
}

extern "C" {

static uint64_t frame_bytes(const rh::Shape& s) {
    uint64_t ext = (uint64_t)s.extW * s.extH;
    return ext * s.ns * 32 + ext * 4 + (uint64_t)s.width() * s.height() * (9 * 4 + 4 + 4) +
           s.buckets.size() * (sizeof(BucketDev) + 4 + 8 + 4 + 64);
}

}  // extern "C"

// Device bytes of one frame of a scheduled scene.
static uint64_t footprint_bytes(const rh::Shape& s, const rh::Schedule& S, const rh_prim* prims, int32_t n) {
    uint64_t bytes = frame_bytes(s) + S.grids.size() * sizeof(GridDev) + S.paramFloats * 4 +
                     (uint64_t)n * sizeof(PrimDev) + S.candGrid.size() * (4 + 4 + 17 * 4) +
                     S.micropolygons * (RH_MPREC_FLOATS * 4 + 16) + S.capStart.back() * 4 +
                     (S.motion ? S.micropolygons * RH_MPREC_MB_FLOATS * 4 : 0);
    bool transparent = false;
    for (int32_t i = 0; i < n; ++i)
        if (rh::hits_listed(s.o, prims[i])) transparent = true;
    if (transparent) {
        // hit arena (24 B an entry), combine list and front depths
        const uint64_t nsamp = (uint64_t)s.extW * s.extH * s.ns;
        bytes += arena_guess(s, S, prims) * 24 + nsamp * 20;
    }
    // device buffers are allocated with 1/8 slack (DBuf::ensure)
    return bytes + bytes / 8 + 64 * 256;
}

// Used by the serving loop: the split/dice schedule of a job, computed once
// when it is submitted, with its footprint; and installing it in a context
// (which, when surfaces may be occlusion culled, computes the culling
// schedule there on the device).
bool rh_internal_schedule(const rh_options& opt, const rh_prim* prims, int32_t n, int threads, rh::Schedule& S,
                          rh::ForestPtr& forest, uint64_t& bytes, double& ms, std::string& err) {
    auto t0 = std::chrono::steady_clock::now();
    rh::Shape s;
    if (!rh::make_shape(opt, s, err)) return false;
    if (!rh::build_forest(s, prims, n, forest, err, threads)) return false;
    bool cullable = false;
    for (int32_t i = 0; i < n && !cullable && rh::cullable_hits(opt); ++i) cullable = !prims[i].no_cull_hidden;
    if (!(cullable ? rh::depth_schedule(s, prims, n, *forest, S, err)
                   : rh::simulate(s, prims, n, *forest, nullptr, nullptr, S, err)))
        return false;
    bytes = footprint_bytes(s, S, prims, n);
    if (cullable) {
        // the tests' scratch: depth pyramids, surface bounds and ranges, pair results
        rh::CullPairs pairs;
        rh::cull_pairs(s, *forest, pairs);
        std::vector<CullTreeDev> trees;
        const uint64_t pyr = rh::cull_trees(s, trees);
        const uint64_t extra = pyr * 4 + trees.size() * sizeof(CullTreeDev) + (pairs.start.size() - 1) * 40 +
                               pairs.start.back() * 4;
        bytes += extra + extra / 8;
    }
    ms = ms_since(t0);
    return true;
}

rh_status rh_internal_post_scheduled(rh_ctx* c, const rh_prim* prims, int32_t n, rh::Schedule&& S,
                                     rh::ForestPtr&& forest, double ms) {
    c->prims.assign(prims, prims + n);
    c->dropGraph();
    const auto t0 = std::chrono::steady_clock::now();
    rh_status r = install_schedule(c, std::move(S), *forest, t0);
    c->stats.ms_schedule += ms;
    if (r == RH_OK) c->rendered = false;
    return r;
}

extern "C" {

rh_status rh_estimate(const rh_options* opt, const rh_prim* prims, int32_t n, rh_footprint* out) {
    if (!opt || !out || n < 0 || (n > 0 && !prims)) return fail(RH_ERR_INVALID, "bad arguments");
    std::string err;
    rh::Schedule S;
    rh::ForestPtr f;
    int threads = (int)std::thread::hardware_concurrency();
    if (threads > 16) threads = 16;
    uint64_t bytes = 0;
    double ms = 0.0;
    if (!rh_internal_schedule(*opt, prims, n, threads, S, f, bytes, ms, err)) return fail(RH_ERR_INVALID, err);
    // (the culling-free schedule: occlusion culling only removes work)
    out->surfaces = S.surfaces;
    out->grids = S.grids.size();
    out->micropolygons = S.micropolygons;
    out->vertices = S.vertices;
    out->list_capacity = S.capStart.back();
    out->device_bytes = bytes;
    return RH_OK;
}

rh_status rh_sample_positions(const rh_options* opt, float* xy, float* dither) {
    if (!opt) return fail(RH_ERR_INVALID, "null options");
    rh::Shape s;
    std::string err;
    if (!rh::make_shape(*opt, s, err)) return fail(RH_ERR_INVALID, err);
    rh::Plan p;
    rh::make_plan(s, p);
    if (xy) {
        for (int y = 0; y < s.extH; ++y)
            for (int x = 0; x < s.extW; ++x) {
                const uint8_t* q = &p.pixPattern[((size_t)y * s.extW + x) * 4];
                for (int i = 0; i < s.ns; ++i) {
                    size_t o = (((size_t)y * s.extW + x) * s.ns + i) * 2;
                    xy[o] = (float)(x + s.extX0) + p.pos2d[2 * ((size_t)q[0] * s.ns + i)];
                    xy[o + 1] = (float)(y + s.extY0) + p.pos2d[2 * ((size_t)q[0] * s.ns + i) + 1];
                }
            }
    }
    if (dither) std::memcpy(dither, p.dither.data(), p.dither.size() * sizeof(float));
    return RH_OK;
}

}  // extern "C"
