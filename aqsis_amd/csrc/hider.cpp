// C ABI of the hider (include/aqsis_hider.h): context, device memory, the
// per-frame launch sequence and read-back.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/aqsis_hider.h"
#include "devtypes.h"
#include "front.h"

extern "C" hipError_t rh_launch_dice(hipStream_t, const FrameDev*, const GridDev*, int, const float*, float4*, float*,
                                     float4*);
extern "C" hipError_t rh_launch_bin(hipStream_t, const FrameDev*, const BucketDev*, int, const TaskDev*,
                                    const GridDev*, const uint32_t*, const uint32_t*, const unsigned long long*,
                                    const float4*, uint32_t*, uint32_t*, uint32_t*, unsigned long long*);
extern "C" hipError_t rh_launch_sample(hipStream_t, const FrameDev*, const BucketDev*, const TaskDev*, int,
                                       const unsigned long long*, const uint32_t*, const uint32_t*, const float*,
                                       const GridDev*, const PrimDev*, const float*, const float*, const uchar4*,
                                       float4*, uint32_t*);
extern "C" hipError_t rh_launch_reduce_counters(hipStream_t, const uint32_t*, int, const uint32_t*, int,
                                                unsigned long long*);
extern "C" size_t rh_filter_tile_lds(int, int, int);
extern "C" hipError_t rh_launch_box_masks(hipStream_t, const FrameDev*, size_t, const uchar4*, const float*,
                                          uint16_t*);
extern "C" hipError_t rh_launch_filter(hipStream_t, const FrameDev*, int, int, int, int, int, const float4*,
                                       const uint16_t*, const uchar4*, const float*, const float*, const float*,
                                       float*, uchar4*);

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

struct DBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes == 0) bytes = 16;
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t n = bytes + bytes / 8 + 256;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) cap = n;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
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

struct rh_ctx {
    rh::Shape shape;
    rh::Plan plan;
    std::vector<rh_prim> prims;
    rh::Schedule sched;
    bool scheduled = false;
    bool rendered = false;
    hipStream_t st = nullptr;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    // frame shape
    DBuf dFrame, dBuckets, dTasks, dPos2d, dPos1d, dPix, dDither, dFilter, dBoxMask;
    bool tiledFilter = false;
    // schedule
    DBuf dGrids, dParams, dPrims, dCandStart, dCandGrid, dCapStart;
    // per frame
    DBuf dVtx, dMprec, dMpbound, dLists, dListCount, dSamples, dChannels, dRgba, dCounters, dTaskCounts, dBucketRefs;
    rh_stats stats;
    ~rh_ctx() {
        DBuf* all[] = {&dFrame, &dBuckets, &dTasks, &dPos2d, &dPos1d, &dPix, &dDither, &dFilter, &dBoxMask,
                       &dGrids, &dParams,
                       &dPrims, &dCandStart, &dCandGrid, &dCapStart, &dVtx, &dMprec, &dMpbound, &dLists,
                       &dListCount, &dSamples, &dChannels, &dRgba, &dCounters, &dTaskCounts, &dBucketRefs};
        for (DBuf* b : all) b->release();
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
}

const char* rh_last_error(void) { return g_err.c_str(); }

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
    rh::make_plan(c->shape, c->plan);
    const rh::Shape& s = c->shape;
    hipError_t e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
    for (int i = 0; i < 5 && e == hipSuccess; ++i) e = hipEventCreate(&c->ev[i]);
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

    std::vector<BucketDev> bk;
    for (size_t k = 0; k < s.buckets.size(); ++k) {
        const auto& b = s.buckets[k];
        BucketDev d;
        d.x = b.x; d.y = b.y; d.w = b.w; d.h = b.h;
        d.sx0 = b.sx0; d.sy0 = b.sy0; d.sx1 = b.sx1; d.sy1 = b.sy1;
        d.bx = s.bx0 + (int)(k % s.nbr); d.by = s.by0 + (int)(k / s.nbr);
        d.task0 = s.taskStart[k];
        d.ntask = s.taskStart[k + 1] - s.taskStart[k];
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
        (e = c->dTaskCounts.ensure(s.tasks.size() * 5 * sizeof(uint32_t))) != hipSuccess ||
        (e = c->dBucketRefs.ensure(s.buckets.size() * sizeof(uint32_t))) != hipSuccess ||
        (e = upload(c->dCounters, zero, c->st)) != hipSuccess) {
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
    if ((e = hipStreamSynchronize(c->st)) != hipSuccess) {
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
    return RH_OK;
}

rh_status rh_clear_surfaces(rh_ctx* c) {
    if (!c) return fail(RH_ERR_INVALID, "null context");
    c->prims.clear();
    c->scheduled = false;
    c->rendered = false;
    return RH_OK;
}

static rh_status schedule(rh_ctx* c) {
    auto t0 = std::chrono::steady_clock::now();
    std::string err;
    int threads = (int)std::thread::hardware_concurrency();
    if (threads > 16) threads = 16;
    if (!rh::make_schedule(c->shape, c->prims.data(), (int)c->prims.size(), c->sched, err, threads))
        return fail(RH_ERR_INVALID, err);
    c->stats.ms_schedule = ms_since(t0);
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
        p[i].pad[0] = p[i].pad[1] = 0;
    }
    std::vector<unsigned long long> cap(S.capStart.begin(), S.capStart.end());
    HIP_TRY(upload(c->dGrids, g, c->st));
    HIP_TRY(upload(c->dParams, S.params, c->st));
    HIP_TRY(upload(c->dPrims, p, c->st));
    HIP_TRY(upload(c->dCandStart, S.candStart, c->st));
    HIP_TRY(upload(c->dCandGrid, S.candGrid, c->st));
    HIP_TRY(upload(c->dCapStart, cap, c->st));
    HIP_TRY(c->dVtx.ensure(S.vertices * sizeof(float4)));
    HIP_TRY(c->dMprec.ensure(S.micropolygons * RH_MPREC_FLOATS * sizeof(float)));
    HIP_TRY(c->dMpbound.ensure(S.micropolygons * sizeof(float4)));
    HIP_TRY(c->dLists.ensure((size_t)S.capStart.back() * sizeof(uint32_t)));
    HIP_TRY(hipStreamSynchronize(c->st));
    c->stats.ms_upload = ms_since(t1);
    c->scheduled = true;
    return RH_OK;
}

rh_status rh_render_device(rh_ctx* c) {
    if (!c) return fail(RH_ERR_INVALID, "null context");
    if (!c->scheduled) return fail(RH_ERR_STATE, "rh_render must run before rh_render_device");
    const rh::Schedule& S = c->sched;
    const rh::Shape& s = c->shape;
    const FrameDev* f = c->dFrame.as<FrameDev>();
    const int nb = (int)s.buckets.size();
    HIP_TRY(hipMemsetAsync(c->dCounters.p, 0, RH_NCOUNTERS * sizeof(unsigned long long), c->st));
    HIP_TRY(hipEventRecord(c->ev[0], c->st));
    HIP_TRY(rh_launch_dice(c->st, f, c->dGrids.as<GridDev>(), (int)S.grids.size(), c->dParams.as<float>(),
                           c->dVtx.as<float4>(), c->dMprec.as<float>(), c->dMpbound.as<float4>()));
    HIP_TRY(hipEventRecord(c->ev[1], c->st));
    HIP_TRY(rh_launch_bin(c->st, f, c->dBuckets.as<BucketDev>(), nb, c->dTasks.as<TaskDev>(),
                          c->dGrids.as<GridDev>(), c->dCandStart.as<uint32_t>(), c->dCandGrid.as<uint32_t>(),
                          c->dCapStart.as<unsigned long long>(), c->dMpbound.as<float4>(),
                          c->dLists.as<uint32_t>(), c->dListCount.as<uint32_t>(), c->dBucketRefs.as<uint32_t>(),
                          c->dCounters.as<unsigned long long>()));
    HIP_TRY(hipEventRecord(c->ev[2], c->st));
    HIP_TRY(rh_launch_sample(c->st, f, c->dBuckets.as<BucketDev>(), c->dTasks.as<TaskDev>(), (int)s.tasks.size(),
                             c->dCapStart.as<unsigned long long>(), c->dListCount.as<uint32_t>(),
                             c->dLists.as<uint32_t>(), c->dMprec.as<float>(), c->dGrids.as<GridDev>(),
                             c->dPrims.as<PrimDev>(), c->dPos2d.as<float>(), c->dPos1d.as<float>(),
                             c->dPix.as<uchar4>(), c->dSamples.as<float4>(), c->dTaskCounts.as<uint32_t>()));
    HIP_TRY(hipEventRecord(c->ev[3], c->st));
    HIP_TRY(rh_launch_reduce_counters(c->st, c->dTaskCounts.as<uint32_t>(), (int)s.tasks.size(),
                                      c->dBucketRefs.as<uint32_t>(), nb, c->dCounters.as<unsigned long long>()));
    HIP_TRY(rh_launch_filter(c->st, f, s.width(), s.height(), s.ns, s.shiftX, s.shiftY, c->dSamples.as<float4>(),
                             c->tiledFilter ? c->dBoxMask.as<uint16_t>() : nullptr, c->dPix.as<uchar4>(),
                             c->dPos2d.as<float>(), c->dFilter.as<float>(), c->dDither.as<float>(),
                             c->dChannels.as<float>(), c->dRgba.as<uchar4>()));
    HIP_TRY(hipEventRecord(c->ev[4], c->st));
    HIP_TRY(hipEventSynchronize(c->ev[4]));
    unsigned long long cnt[RH_NCOUNTERS];
    HIP_TRY(hipMemcpy(cnt, c->dCounters.p, sizeof(cnt), hipMemcpyDeviceToHost));
    float t[4];
    for (int i = 0; i < 4; ++i) HIP_TRY(hipEventElapsedTime(&t[i], c->ev[i], c->ev[i + 1]));
    c->stats.spl_count = cnt[0];
    c->stats.spl_bound_hits = cnt[1];
    c->stats.spl_hits = cnt[2];
    c->stats.inside = cnt[3];
    c->stats.inside_nonlinear = cnt[4];
    c->stats.grids = S.grids.size();
    c->stats.micropolygons = S.micropolygons;
    c->stats.surfaces = S.surfaces;
    c->stats.occlusion_culled = 0;
    c->stats.ms_dice = t[0];
    c->stats.ms_bin = t[1];
    c->stats.ms_sample = t[2];
    c->stats.ms_filter = t[3];
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

rh_status rh_read_samples(rh_ctx* c, uint32_t* mp_id, float* depth, float* rgbaz) {
    if (!c) return fail(RH_ERR_INVALID, "null context");
    if (!c->rendered) return fail(RH_ERR_STATE, "no rendered frame");
    const size_t n = (size_t)c->shape.extW * c->shape.extH * c->shape.ns;
    std::vector<float> buf(n * 8);
    HIP_TRY(hipMemcpy(buf.data(), c->dSamples.p, n * 32, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) {
        const float* r = &buf[i * 8];
        uint32_t id;
        std::memcpy(&id, &r[7], 4);
        if (mp_id) mp_id[i] = id;
        if (depth) depth[i] = r[3];
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

static uint64_t frame_bytes(const rh::Shape& s) {
    uint64_t ext = (uint64_t)s.extW * s.extH;
    return ext * s.ns * 32 + ext * 4 + (uint64_t)s.width() * s.height() * (9 * 4 + 4 + 4) +
           s.buckets.size() * (sizeof(BucketDev) + 4 + 8 + 4 + 64);
}

rh_status rh_estimate(const rh_options* opt, const rh_prim* prims, int32_t n, rh_footprint* out) {
    if (!opt || !out || n < 0 || (n > 0 && !prims)) return fail(RH_ERR_INVALID, "bad arguments");
    rh::Shape s;
    std::string err;
    if (!rh::make_shape(*opt, s, err)) return fail(RH_ERR_INVALID, err);
    rh::Schedule S;
    int threads = (int)std::thread::hardware_concurrency();
    if (threads > 16) threads = 16;
    if (!rh::make_schedule(s, prims, n, S, err, threads)) return fail(RH_ERR_INVALID, err);
    out->surfaces = S.surfaces;
    out->grids = S.grids.size();
    out->micropolygons = S.micropolygons;
    out->vertices = S.vertices;
    out->list_capacity = S.capStart.back();
    out->device_bytes = frame_bytes(s) + S.grids.size() * sizeof(GridDev) + S.params.size() * 4 +
                        (uint64_t)n * sizeof(PrimDev) + S.candGrid.size() * 4 + S.vertices * 16 +
                        S.micropolygons * (RH_MPREC_FLOATS * 4 + 16) + S.capStart.back() * 4;
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
