// Host front end: frame shape, sample plan, split/dice schedule.
// Compile with -O2 -ffp-contract=off (x86-64 SSE, no FMA contraction).
#include "front.h"

#include <algorithm>
#include <functional>
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <pthread.h>
#include <random>
#include <thread>

namespace rh {

// ---------------------------------------------------------------------------
// Host worker pool
// ---------------------------------------------------------------------------
// A process-wide pool of worker threads, started on first use and kept until
// the process ends (never destroyed: its threads wait on a condition).
// Starting and joining threads per parallel loop cost tens of microseconds a
// thread, which made up most of a small scene's schedule.  One loop runs on
// the pool at a time (callers on other threads wait for it); the caller
// takes part in its own loop.
namespace {
struct WorkPool {
    std::mutex runMutex;                  // one loop at a time
    std::mutex m;
    std::condition_variable wake, finished;
    std::vector<std::thread> workers;
    const std::function<void(int)>* fn = nullptr;
    std::atomic<int> next{0};
    int count = 0;
    int wanted = 0;                       // workers taking part in the current loop
    int done = 0;                         // of those, the ones finished
    uint64_t gen = 0;

    void loop(int id) {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(m);
        for (;;) {
            wake.wait(lk, [&] { return gen != seen; });
            seen = gen;
            if (id >= wanted) continue;
            const std::function<void(int)>* f = fn;
            const int cnt = count;
            lk.unlock();
            for (;;) {
                const int i = next.fetch_add(1);
                if (i >= cnt) break;
                (*f)(i);
            }
            lk.lock();
            if (++done == wanted) finished.notify_all();
        }
    }

    void run(int n, int threads, const std::function<void(int)>& f) {
        std::lock_guard<std::mutex> one(runMutex);
        const int helpers = std::min(threads, n) - 1;
        {
            std::unique_lock<std::mutex> lk(m);
            while ((int)workers.size() < helpers) {
                const int id = (int)workers.size();
                workers.emplace_back([this, id] { loop(id); });
                workers.back().detach();
            }
            fn = &f;
            count = n;
            next = 0;
            wanted = helpers;
            done = 0;
            ++gen;
        }
        wake.notify_all();
        for (;;) {
            const int i = next.fetch_add(1);
            if (i >= n) break;
            f(i);
        }
        std::unique_lock<std::mutex> lk(m);
        finished.wait(lk, [&] { return done == wanted; });
        fn = nullptr;
    }
};

WorkPool* g_pool = nullptr;

// A forked child has none of the parent's workers: it starts a pool of its own.
void pool_after_fork() { g_pool = new WorkPool(); }

WorkPool& work_pool() {
    static const bool registered = [] {
        g_pool = new WorkPool();              // kept until the process ends
        pthread_atfork(nullptr, nullptr, pool_after_fork);
        return true;
    }();
    (void)registered;
    return *g_pool;
}
}  // namespace

// Run fn(i) for i in [0, n) on up to `threads` threads (the caller and the
// pool's workers), in no particular order.
static void run_parallel(int n, int threads, const std::function<void(int)>& fn) {
    if (n <= 0) return;
    if (std::min(threads, n) <= 1) {
        for (int i = 0; i < n; ++i) fn(i);
        return;
    }
    work_pool().run(n, threads, fn);
}

// ---------------------------------------------------------------------------
// Frame shape
// ---------------------------------------------------------------------------
static float filter_value(int f, float x, float y, float xw, float yw);

bool make_shape(const rh_options& o, Shape& s, std::string& err) {
    if (o.xres <= 0 || o.yres <= 0 || o.xres > 32768 || o.yres > 32768) { err = "resolution out of range"; return false; }
    if (o.xsamples < 1 || o.ysamples < 1 || o.xsamples * o.ysamples > 64) {
        err = "PixelSamples must be >= 1 each with at most 64 samples per pixel";
        return false;
    }
    if (!(o.filter_xwidth > 0) || !(o.filter_ywidth > 0) || o.filter_xwidth > 8 || o.filter_ywidth > 8) {
        err = "filter widths must be in (0, 8]";
        return false;
    }
    if (o.filter < RH_FILTER_BOX || o.filter > RH_FILTER_MITCHELL) { err = "unknown pixel filter"; return false; }
    if (o.bucket_x < 4 || o.bucket_y < 4 || o.bucket_x > 64 || o.bucket_y > 64) { err = "bucket size must be in [4, 64]"; return false; }
    if (!(o.fov > 0 && o.fov < 180)) { err = "fov must be in (0, 180)"; return false; }
    if (o.num_displays < 0 || o.num_displays > 16) { err = "num_displays out of range"; return false; }
    if (!(o.pixel_aspect > 0)) { err = "pixel aspect must be positive"; return false; }
    if (o.depth_filter < RH_DEPTH_MIN || o.depth_filter > RH_DEPTH_AVERAGE) { err = "unknown depth filter"; return false; }
    if (o.dof[0] < FLT_MAX && !(o.dof[0] > 0 && o.dof[1] > 0 && o.dof[2] > 0)) {
        err = "DepthOfField fstop, focal length and focal distance must be positive";
        return false;
    }
    s.o = o;
    // CqRenderer::initialiseCropWindow (renderer.cpp:1619-1626)
    s.cropX0 = (int)clampv<long>(ceil_l((float)o.xres * o.crop[0]), 0, o.xres);
    s.cropX1 = (int)clampv<long>(ceil_l((float)o.xres * o.crop[1]), 0, o.xres);
    s.cropY0 = (int)clampv<long>(ceil_l((float)o.yres * o.crop[2]), 0, o.yres);
    s.cropY1 = (int)clampv<long>(ceil_l((float)o.yres * o.crop[3]), 0, o.yres);
    if (s.cropX1 <= s.cropX0 || s.cropY1 <= s.cropY0) { err = "empty crop window"; return false; }
    s.shiftX = (int)floor_l(o.filter_xwidth / 2.0f);
    s.shiftY = (int)floor_l(o.filter_ywidth / 2.0f);
    s.xfwo2 = std::ceil(o.filter_xwidth) * 0.5f;
    s.yfwo2 = std::ceil(o.filter_ywidth) * 0.5f;
    if (s.shiftX > 3 || s.shiftY > 3) { err = "filter footprint too large"; return false; }
    s.nx = o.xsamples; s.ny = o.ysamples; s.ns = s.nx * s.ny;
    // CqImageBuffer::SetImage (imagebuffer.cpp:174-228)
    s.nbx = (o.xres - 1) / o.bucket_x + 1;
    s.nby = (o.yres - 1) / o.bucket_y + 1;
    s.bx0 = s.cropX0 / o.bucket_x; s.by0 = s.cropY0 / o.bucket_y;
    s.bx1 = (s.cropX1 - 1) / o.bucket_x + 1; s.by1 = (s.cropY1 - 1) / o.bucket_y + 1;
    s.nbr = s.bx1 - s.bx0; s.nbrY = s.by1 - s.by0;
    s.extX0 = s.cropX0 - s.shiftX; s.extY0 = s.cropY0 - s.shiftY;
    s.extW = s.width() + 2 * s.shiftX; s.extH = s.height() + 2 * s.shiftY;

    // Camera (ri.cpp:586-617; options.cpp:40-160), perspective projection.
    float par = o.pixel_aspect;
    float far_ = ((float)o.xres * par) / (float)o.yres;
    float l, r, t, b;
    if (far_ >= 1.0) { l = -far_; r = far_; t = 1; b = -1; }
    else { l = -1; r = 1; t = (float)(1.0 / far_); b = (float)(-1.0 / far_); }
    float n = o.clip_near, f = o.clip_far;
    float sfov = (float)std::tan((double)(deg2rad(o.fov) / 2.0f));
    float a = f / (f - n);
    m44 c2i = m_identity();
    c2i.id = false;
    c2i.e[0][0] = 1.0f / sfov; c2i.e[1][1] = 1.0f / sfov; c2i.e[2][2] = a;
    c2i.e[3][2] = -n * a; c2i.e[2][3] = 1.0f; c2i.e[3][3] = 0.0f;
    m44 i2s = m_identity();
    i2s.id = false;
    i2s.e[0][0] = 2.0f / (r - l); i2s.e[1][1] = 2.0f / (t - b);
    i2s.e[3][0] = (-(r + l)) / (r - l); i2s.e[3][1] = (-(t + b)) / (t - b);
    m44 c2s = m_mul(m_mul(i2s, m_identity()), c2i);
    float frameX = (far_ >= 1.0) ? (float)o.xres : ((float)o.yres * far_) / par;
    float frameY = (far_ < 1.0) ? (float)o.yres : ((float)o.xres * par) / far_;
    m44 s2ndc = m_mul(m_mul(m_diag(0.5f, 0.5f, 1.0f), m_trans(1.0f, 1.0f, 0.0f)), m_diag(1.0f, -1.0f, 1.0f));
    m44 raster = m_mul(m_mul(m_mul(m_diag(frameX, frameY, 1.0f), s2ndc), c2s), m_identity());
    s.camToRaster = m_mul(raster, m_identity());
    s.diceM = s.camToRaster;
    s.diceM.e[0][2] = s.diceM.e[1][2] = s.diceM.e[2][2] = s.diceM.e[3][2] = 0;

    // Depth of field.  The CoC scale maps (1,1,-1) and (0,0,-1) through
    // matVSpaceToSpace("camera", "raster"): the same product with its
    // translation and projective column cleared (renderer.cpp:897-957,
    // options.cpp:160-171).  The multiplier and inverse focal distance are
    // computed in double and stored as floats (renderer.h:368-377).
    {
        m44 dm = s.camToRaster;
        dm.e[3][0] = dm.e[3][1] = dm.e[3][2] = dm.e[0][3] = dm.e[1][3] = dm.e[2][3] = 0.0f;
        dm.e[3][3] = 1.0f;
        const f3 de = m_point(dm, mk(1, 1, -1)), dc = m_point(dm, mk(0, 0, -1));
        s.dofScaleX = std::fabs(de.x - dc.x);
        s.dofScaleY = std::fabs(de.y - dc.y);
        s.dof = o.dof[0] < FLT_MAX;
        s.dofCells.clear();
        if (s.dof) {
            const float fstop = o.dof[0], focalLength = o.dof[1], focalDistance = o.dof[2];
            const float lensDiameter = focalLength / fstop;
            s.dofMult = (float)(0.5 * lensDiameter * focalDistance / (focalDistance + lensDiameter));
            s.dofOneOverF = (float)(1.0 / focalDistance);
            // CalculateDofBounds (bucketprocessor.cpp:858-911): each cell of
            // [-1,1]^2 bounded by its projected corners, or by the square's
            // own extent across an axis it straddles.
            s.dofCells.assign((size_t)s.ns * 4, 0.0f);
            const float dx = (float)(2.0 / s.nx), dy = (float)(2.0 / s.ny);
            float minX = -1.0f, minY = -1.0f;
            for (int j = 0, which = 0; j < s.ny; ++j) {
                for (int i = 0; i < s.nx; ++i, ++which) {
                    float tl[2], tr[2], bl[2], br[2];
                    project_to_circle(minX, minY, tl[0], tl[1]);
                    project_to_circle(minX + dx, minY, tr[0], tr[1]);
                    project_to_circle(minX, minY + dy, bl[0], bl[1]);
                    project_to_circle(minX + dx, minY + dy, br[0], br[1]);
                    if ((tl[1] > 0.0f && bl[1] < 0.0f) || (tl[1] < 0.0f && bl[1] > 0.0f)) {
                        tl[0] = minX; bl[0] = minX; tr[0] = minX + dx; br[0] = minX + dx;
                    }
                    if ((tl[0] > 0.0f && tr[0] < 0.0f) || (tl[0] < 0.0f && tr[0] > 0.0f)) {
                        tl[1] = minY; bl[1] = minY + dy; tr[1] = minY; br[1] = minY + dy;
                    }
                    float* c = &s.dofCells[4 * (size_t)which];
                    c[0] = tl[0]; c[1] = tl[1]; c[2] = tl[0]; c[3] = tl[1];
                    for (const float* q : {tr, bl, br}) {
                        c[2] = mx(c[2], q[0]); c[3] = mx(c[3], q[1]);
                        c[0] = mn(c[0], q[0]); c[1] = mn(c[1], q[1]);
                    }
                    minX += dx;
                }
                minX = -1.0f;
                minY += dy;
            }
        }
    }

    // Filter weights at subpixel centres (bucketprocessor.cpp:811-856)
    const int xm = s.shiftX, ym = s.shiftY;
    s.filterTab.assign((size_t)(2 * xm + 1) * (2 * ym + 1) * s.ns, 0.0f);
    float cw = std::ceil(o.filter_xwidth), ch = std::ceil(o.filter_ywidth);
    for (int py = -ym; py <= ym; ++py)
        for (int px = -xm; px <= xm; ++px) {
            size_t k = ((size_t)(py + ym) * (2 * xm + 1) + (px + xm)) * s.ns;
            for (int sy = 0; sy < s.ny; ++sy)
                for (int sx = 0; sx < s.nx; ++sx, ++k) {
                    float fx = (sx + 0.5f) / s.nx + px - 0.5f;
                    float fy = (sy + 0.5f) / s.ny + py - 0.5f;
                    float w = 0;
                    if (fx >= -s.xfwo2 && fy >= -s.yfwo2 && fx <= s.xfwo2 && fy <= s.yfwo2)
                        w = filter_value(o.filter, fx, fy, cw, ch);
                    s.filterTab[k] = w;
                }
        }

    // Buckets and the sample region each one owns (bucketprocessor.cpp:95-137):
    // display rect +- shift, clipped to the crop window +- shift, minus the
    // 2*shift bands handed over from the left and upper neighbours.
    s.buckets.clear();
    for (int by = s.by0; by < s.by1; ++by)
        for (int bx = s.bx0; bx < s.bx1; ++bx) {
            Shape::BucketInfo bi;
            bi.x = bx * o.bucket_x; bi.y = by * o.bucket_y;
            bi.w = std::min(o.xres - bi.x, o.bucket_x);
            bi.h = std::min(o.yres - bi.y, o.bucket_y);
            int sx0 = bi.x - s.shiftX, sy0 = bi.y - s.shiftY;
            int sx1 = bi.x + bi.w + s.shiftX, sy1 = bi.y + bi.h + s.shiftY;
            sx0 = std::max(sx0, s.cropX0 - s.shiftX);
            sy0 = std::max(sy0, s.cropY0 - s.shiftY);
            sx1 = std::min(sx1, s.cropX1 + s.shiftX);
            sy1 = std::min(sy1, s.cropY1 + s.shiftY);
            if (bx > s.bx0) sx0 += 2 * s.shiftX;
            if (by > s.by0) sy0 += 2 * s.shiftY;
            bi.sx0 = sx0; bi.sy0 = sy0; bi.sx1 = sx1; bi.sy1 = sy1;
            s.buckets.push_back(bi);
        }

    // Wave blocks: 64 samples as a near-square power-of-two pixel block, a
    // row of pixels when 64 / ns is not a power of two, one pixel when
    // ns >= 64.
    int ppw = 64 / s.ns;
    if (ppw < 1) ppw = 1;
    if ((ppw & (ppw - 1)) == 0) {
        int lg = 0;
        while ((1 << lg) < ppw) ++lg;
        s.blockW = 1 << ((lg + 1) / 2);
        s.blockH = ppw / s.blockW;
    } else {
        s.blockW = ppw;
        s.blockH = 1;
    }
    // Tasks: split each owned region into at most 4 x 4 near-equal parts of
    // at least 4 blocks per side.
    const int ntx = std::max(1, std::min(4, o.bucket_x / (4 * s.blockW)));
    const int nty = std::max(1, std::min(4, o.bucket_y / (4 * s.blockH)));
    s.tasks.clear();
    s.taskStart.assign(1, 0);
    for (size_t k = 0; k < s.buckets.size(); ++k) {
        const auto& b = s.buckets[k];
        const int ow = b.sx1 - b.sx0, oh = b.sy1 - b.sy0;
        if (ow > 0 && oh > 0) {
            const int tx = std::min(ntx, ow), ty = std::min(nty, oh);
            for (int j = 0; j < ty; ++j)
                for (int i = 0; i < tx; ++i) {
                    Shape::TaskInfo t;
                    t.bucket = (int)k;
                    t.x0 = b.sx0 + (i * ow) / tx; t.x1 = b.sx0 + ((i + 1) * ow) / tx;
                    t.y0 = b.sy0 + (j * oh) / ty; t.y1 = b.sy0 + ((j + 1) * oh) / ty;
                    s.tasks.push_back(t);
                }
        }
        s.taskStart.push_back((int)s.tasks.size());
    }
    return true;
}

void project_to_circle(float x, float y, float& ox, float& oy) {
    float r;   // CqVector2D::Magnitude
    if (y == 0.0f) r = std::fabs(x);
    else if (x == 0.0f) r = std::fabs(y);
    else r = std::sqrt((x * x) + (y * y));
    if (r == 0.0f) { ox = oy = 0.0f; return; }
    const float adj = mx(std::fabs(x), std::fabs(y)) / r;
    ox = adj * x;
    oy = adj * y;
}

// RiXxxFilter functions (libs/core/filters.cpp:72-342)
static float filter_value(int f, float x, float y, float xw, float yw) {
    const double PI = 3.14159265358979323846;
    switch (f) {
        case RH_FILTER_BOX: {
            double a = (std::fabs(x) <= xw / 2.0) ? 1.0 : 0.0, b = (std::fabs(y) <= yw / 2.0) ? 1.0 : 0.0;
            return (float)(a < b ? a : b);
        }
        case RH_FILTER_TRIANGLE: {
            float hx = (float)(xw / 2.0), hy = (float)(yw / 2.0), ax = std::fabs(x), ay = std::fabs(y);
            double a = ax <= hx ? (double)((hx - ax) / hx) : 0.0, b = ay <= hy ? (double)((hy - ay) / hy) : 0.0;
            return (float)(a < b ? a : b);
        }
        case RH_FILTER_CATMULLROM: {
            float r2 = (x * x + y * y);
            float r = std::sqrt(r2);
            return (float)((r >= 2.0) ? 0.0 : (r < 1.0) ? (3.0 * r * r2 - 5.0 * r2 + 2.0) : (-r * r2 + 5.0 * r2 - 8.0 * r + 4.0));
        }
        case RH_FILTER_GAUSSIAN: {
            x /= xw; y /= yw;
            return (float)std::exp(-8.0 * (double)(x * x + y * y));
        }
        case RH_FILTER_SINC: {
            if (x != 0.0) { x = (float)(x * PI); x = (float)(std::cos(0.5 * x / xw) * std::sin((double)x) / x); }
            else x = 1.0f;
            if (y != 0.0) { y = (float)(y * PI); y = (float)(std::cos(0.5 * y / yw) * std::sin((double)y) / y); }
            else y = 1.0f;
            return x * y;
        }
        case RH_FILTER_DISK: {
            double xx = x * x, yy = y * y;
            xw *= 0.5; yw *= 0.5;
            double d = xx / (xw * xw) + yy / (yw * yw);
            return d < 1.0 ? 1.0f : 0.0f;
        }
        case RH_FILTER_BESSEL: {
            double xx = x * x, yy = y * y;
            xw *= 0.5; yw *= 0.5;
            double w = xx / (xw * xw) + yy / (yw * yw);
            if (w < 1.0) {
                double d = std::sqrt(xx + yy);
                if (d != 0.0) { w = std::cos(0.5 * PI * std::sqrt(w)); return (float)(w * 2 * j1(PI * d) / d); }
                return (float)PI;
            }
            return 0.0f;
        }
        case RH_FILTER_MITCHELL: {
            const float B = 1 / 3.0f, C = 1 / 3.0f;
            float ix = 1.0f / xw, iy = 1.0f / yw;
            auto ev = [&](float t) -> float {
                t = std::fabs(2.f * t);
                if (t > 1.f)
                    return ((-B - 6 * C) * t * t * t + (6 * B + 30 * C) * t * t + (-12 * B - 48 * C) * t + (8 * B + 24 * C)) * (1.f / 6.f);
                return ((12 - 9 * B - 6 * C) * t * t * t + (-18 + 12 * B + 6 * C) * t * t + (6 - 2 * B)) * (1.f / 6.f);
            };
            return ev(x * ix) * ev(y * iy);
        }
    }
    return 0.0f;
}

// ---------------------------------------------------------------------------
// Sample plan
// ---------------------------------------------------------------------------
// aqsis's CqRandom is one MT19937 stream (libs/math/random.cpp); std::mt19937
// has the same seeding and output.  RandomFloat = u32 * (1/4294967424) in
// double, narrowed to float; RandomInt(n) = lfloor((double)(n * RandomFloat())).
namespace {
struct Stream {
    std::mt19937 g;
    float f() { return (float)((double)g() * (1.0 / 4294967424.0)); }
    float f(float r) { return r * f(); }
    int i(unsigned n) { double v = f((float)n); return (int)floor_l(v); }
};
}  // namespace

void make_plan(const Shape& s, Plan& p) {
    const int ns = s.ns, nx = s.nx, ny = s.ny;
    Stream st;
    st.g.seed(545);  // 'a'+'q'+'s'+'i'+'s' at WorldBegin (ri.cpp:660)
    if (s.o.jitter) {
        // CqMultiJitteredSampler (multijitter.h:90-101, multijitter.cpp:83-202)
        p.cache = 250;
        p.pos2d.assign((size_t)250 * ns * 2, 0.0f);
        p.pos1d.assign((size_t)250 * ns, 0.0f);
        std::vector<int> shuffled((size_t)250 * ns);
        std::vector<int> cell(2 * ns);
        for (int pat = 0; pat < 250; ++pat) {
            const size_t off = (size_t)pat * ns;
            if (nx == 1 && ny == 1) {
                float ry = st.f(), rx = st.f();  // CqVector2D(f(), f()): g++ evaluates y first
                p.pos2d[2 * off] = rx; p.pos2d[2 * off + 1] = ry;
                p.pos1d[off] = st.f();
            } else {
                for (int iy = 0; iy < ny; ++iy)
                    for (int ix = 0; ix < nx; ++ix) {
                        cell[2 * (iy * nx + ix)] = iy;
                        cell[2 * (iy * nx + ix) + 1] = ix;
                    }
                for (int iy = 0; iy < ny; ++iy)
                    for (int ix = nx; ix > 1;) {
                        int k = st.i(ix);
                        --ix;
                        std::swap(cell[2 * (iy * nx + ix) + 1], cell[2 * (iy * nx + k) + 1]);
                    }
                for (int ix = 0; ix < nx; ++ix)
                    for (int iy = ny; iy > 1;) {
                        int k = st.i(iy);
                        --iy;
                        std::swap(cell[2 * (iy * nx + ix)], cell[2 * (k * nx + ix)]);
                    }
                const float h = 1.0f / ny, w = 1.0f / nx, c = 1.0f / ns;
                for (int iy = 0, k = 0; iy < ny; ++iy)
                    for (int ix = 0; ix < nx; ++ix, ++k) {
                        float ry = st.f(), rx = st.f();
                        p.pos2d[2 * (off + k)] = (cell[2 * k] + rx) * c + ix * w;
                        p.pos2d[2 * (off + k) + 1] = (cell[2 * k + 1] + ry) * c + iy * h;
                    }
            }
            float acc = 0, d = 1.0f / ns, r1 = st.f(d);
            for (int k = 0; k < ns; ++k) { p.pos1d[off + k] = acc + r1; acc += d; }
            for (int k = 0; k < ns; ++k) shuffled[off + k] = k;
            for (int j = ns; j > 1;) {
                int k = st.i(j);
                --j;
                std::swap(shuffled[off + j], shuffled[off + k]);
            }
        }
        st.g.seed(19);
        p.shuffled.assign(shuffled.begin(), shuffled.end());
    } else {
        // CqGridSampler (grid.cpp:37-63): no random draws at all.
        p.cache = 1;
        p.pos2d.assign((size_t)ns * 2, 0.0f);
        p.pos1d.assign(ns, 0.0f);
        float xs = (float)(1.0 / nx), ys = (float)(1.0 / ny);
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                p.pos2d[2 * (j * nx + i)] = (float)(xs * (i + 0.5));
                p.pos2d[2 * (j * nx + i) + 1] = (float)(ys * (j + 0.5));
            }
        float dt = (float)(1 / ns), t = (float)(dt * 0.5);
        for (int i = 0; i < ns; ++i) { p.pos1d[i] = t; t += dt; }
        p.shuffled.resize(ns);
        for (int i = 0; i < ns; ++i) p.shuffled[i] = (uint8_t)i;
    }
    // Lens offsets of each pattern: projectToCircle(-1 + 2 * offset)
    // (imagepixel.cpp:357), offset k going to the sample shuffled[k].
    p.lens.assign(p.pos2d.size(), 0.0f);
    for (size_t k = 0; k + 1 < p.pos2d.size(); k += 2)
        project_to_circle(-1.0f + 2.0f * p.pos2d[k], -1.0f + 2.0f * p.pos2d[k + 1], p.lens[k], p.lens[k + 1]);
    // Bucket loop order: per bucket, 5 draws per sample-region pixel
    // (imagepixel.cpp:338-347), then one draw per display pixel per display
    // request (ddmanager.cpp:1050).
    p.pixPattern.assign((size_t)s.extW * s.extH * 4, 0);
    p.pixDof.assign((size_t)s.extW * s.extH, 0);
    p.dither.assign((size_t)s.width() * s.height(), 0.0f);
    const size_t npix = (size_t)s.width() * s.height();
    p.ditherAll.assign(npix * (size_t)std::max(1, s.o.num_displays), 0.0f);
    for (const auto& b : s.buckets) {
        if (s.o.jitter) {
            for (int y = b.sy0; y < b.sy1; ++y)
                for (int x = b.sx0; x < b.sx1; ++x) {
                    uint8_t* q = &p.pixPattern[((size_t)(y - s.extY0) * s.extW + (x - s.extX0)) * 4];
                    int shuf = st.i(250);
                    int pos = st.i(250);
                    int dof = st.i(250);
                    int tim = st.i(250);
                    int lod = st.i(250);
                    q[0] = (uint8_t)pos; q[1] = (uint8_t)tim; q[2] = (uint8_t)lod; q[3] = (uint8_t)shuf;
                    p.pixDof[(size_t)(y - s.extY0) * s.extW + (x - s.extX0)] = (uint8_t)dof;
                }
        }
        bool outside = (b.x + b.w) <= s.cropX0 || (b.y + b.h) <= s.cropY0 || b.x > s.cropX1 || b.y > s.cropY1;
        if (outside) continue;
        for (int d = 0; d < s.o.num_displays; ++d)
            for (int y = b.y; y < b.y + b.h; ++y)
                for (int x = b.x; x < b.x + b.w; ++x) {
                    float v = st.f();
                    if (x >= s.cropX0 && x < s.cropX1 && y >= s.cropY0 && y < s.cropY1) {
                        const size_t i = (size_t)(y - s.cropY0) * s.width() + (x - s.cropX0);
                        p.ditherAll[(size_t)d * npix + i] = v;
                        if (d == 0) p.dither[i] = v;
                    }
                }
    }
}

// ---------------------------------------------------------------------------
// Split / dice schedule
// ---------------------------------------------------------------------------
namespace {

enum { DIR_U = 0, DIR_V = 1 };

struct Node {
    int type = 0, prim = 0;
    int splitCount = 0, splitDir = DIR_U;
    bool diceableFlag = true;   // m_fDiceable
    bool discard = false;       // m_fDiscard
    bool culled = false;        // rejected by CullSurface
    bool hasBound = false;      // cached raster bound present
    bool leaf = false;          // Diceable() returned true
    int uDice = 0, vDice = 0;
    box3 rb;                    // cached raster bound
    int xminb = 0, yminb = 0, xmaxb = 0, ymaxb = 0;
    int kid0 = -1, nkids = 0;   // children in the same tree
    // geometry (key 0 of a deforming surface)
    float radius = 1, phiMin = 0, phiMax = 0, thetaMin = 0, thetaMax = 360;
    f3 cp[16];
    // A deforming surface (CqDeformingSurface, geometry/surface.h:617-720):
    // key 1's control points are in its tree's cp1 (spheres: key 1's matrix,
    // Ctx::objToCam1).
    bool motion = false;
    float subMin = FLT_MAX;     // least cached minimum depth in the subtree
};

// One primitive's split tree: nodes, parents before children, and for a
// deforming surface key 1's control points, 16 per node.
struct Tree {
    std::vector<Node> n;
    std::vector<f3> cp1;
};

// sinCosGrid (quadrics.cpp:1493-1518)
void sincos_grid(float t0, float t1, int n, float* s, float* c) {
    double pc = std::cos((double)t0), ps = std::sin((double)t0);
    double dt = (double)((t1 - t0) / (n - 1));
    double cd = std::cos(dt), sd = std::sin(dt);
    c[0] = (float)pc; s[0] = (float)ps;
    for (int i = 1; i < n; ++i) {
        double cc = cd * pc - sd * ps, ss = sd * pc + cd * ps;
        c[i] = (float)cc; s[i] = (float)ss;
        pc = cc; ps = ss;
    }
}

// IntersectLine (quadrics.cpp:1527-1541)
bool line_meet(f3 P1, f3 T1, f3 P2, f3 T2, f3& P) {
    f3 v = crs(crs(T1, sub(P1, T2)), T1);
    float vw = dot3(v, T2);
    if ((double)(vw * vw) < 1.0e-07) return false;
    P = add(P2, scale(dot3(sub(P1, P2), v) / vw, T2));
    return true;
}

box3 sphere_camera_bound(const Node& n, const m44& objToCam) {
    const double PI = 3.14159265358979323846;
    // Circle(O=(0,0,0), X=(1,0,0), Y=(0,0,1), r, min phi, max phi) (quadrics.cpp:1360-1400)
    float as = mn(n.phiMin, n.phiMax), ae = mx(n.phiMin, n.phiMax), r = n.radius;
    while (ae < as) ae = (float)(ae + 2 * PI);
    float dth = (ae - as) / 4.0f;
    f3 X = mk(1, 0, 0), Y = mk(0, 0, 1);
    f3 prof[9];
    f3 P0 = add(add(mk(0, 0, 0), scale((float)(r * std::cos((double)as)), X)), scale((float)(r * std::sin((double)as)), Y));
    f3 T0 = add(scale((float)(-std::sin((double)as)), X), scale((float)std::cos((double)as), Y));
    f3 P1 = mk(0, 0, 0);
    prof[0] = P0;
    float ang = as;
    for (int i = 1; i <= 4; ++i) {
        ang += dth;
        f3 P2 = add(add(mk(0, 0, 0), scale((float)(r * std::cos((double)ang)), X)), scale((float)(r * std::sin((double)ang)), Y));
        prof[2 * i] = P2;
        f3 T2 = add(scale((float)(-std::sin((double)ang)), X), scale((float)std::cos((double)ang), Y));
        line_meet(P0, T0, P2, T2, P1);
        prof[2 * i - 1] = P1;
        if (i < 4) { P0 = P2; T0 = T2; }
    }
    m44 rot = m_rot(deg2rad(n.thetaMin), Y);
    for (auto& q : prof) q = m_point(rot, q);
    // RevolveForBound(profile, S=(0,0,0), Tvec=(0,0,1), theta) (quadrics.cpp:1404-1474)
    float theta = deg2rad(n.thetaMax - n.thetaMin);
    if (std::fabs(theta) > 2.0 * PI) theta = theta < 0 ? (float)(-(2.0 * PI)) : (float)(2.0 * PI);
    float dt = theta / 4.0f;
    float cs[5], sn[5];
    for (int i = 1; i <= 4; ++i) {
        float a = dt * (float)i;
        cs[i] = (float)std::cos((double)a);
        sn[i] = (float)std::sin((double)a);
    }
    box3 b = empty_box();
    f3 Tv = mk(0, 0, 1), S = mk(0, 0, 0);
    P1 = mk(0, 0, 0);
    for (int j = 0; j < 9; ++j) {
        f3 pj = prof[j];
        f3 a = sub(pj, S);
        float den = len2(Tv);
        float frac = (den == 0.0) ? 0.0f : dot3(Tv, a) / den;
        f3 O = add(scale(frac, Tv), S);
        f3 Xa = sub(pj, O);
        float rr = std::sqrt(len2(Xa));
        if ((double)rr < 1e-7) { grow(b, O); continue; }
        Xa = normalised(Xa);
        f3 Ya = normalised(crs(Tv, Xa));
        f3 Q0 = pj;
        grow(b, Q0);
        f3 U0 = Ya;
        for (int i = 1; i <= 4; ++i) {
            f3 Q2 = add(add(O, scale(rr * cs[i], Xa)), scale(rr * sn[i], Ya));
            grow(b, Q2);
            f3 U2 = add(scale(-sn[i], Xa), scale(cs[i], Ya));
            line_meet(Q0, U0, Q2, U2, P1);
            grow(b, P1);
            if (i < 4) { Q0 = Q2; U0 = U2; }
        }
    }
    return box_xform(b, objToCam);
}

box3 patch_camera_bound(const Node& n, int k) {
    box3 b;
    b.lo = mk(FLT_MAX, FLT_MAX, FLT_MAX);
    b.hi = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    for (int i = 0; i < k; ++i) {
        f3 v = n.cp[i];
        if (v.x < b.lo.x) b.lo.x = v.x;
        if (v.y < b.lo.y) b.lo.y = v.y;
        if (v.x > b.hi.x) b.hi.x = v.x;
        if (v.y > b.hi.y) b.hi.y = v.y;
        if (v.z < b.lo.z) b.lo.z = v.z;
        if (v.z > b.hi.z) b.hi.z = v.z;
    }
    return b;
}

struct Ctx {
    const Shape* s;
    const rh_prim* prims;
    std::vector<m44> objToCam;   // sphere transforms
    std::vector<m44> objToCam1;  // key 1 of moving spheres
};

// CqSurface::AdjustedShadingRate (geometry/surface.cpp:688-724) without the
// motion factor: with depth of field, ShadingRate times max(1, 0.025 *
// GeometricFocusFactor * MinCoCForBound^2) over the cached raster bound's
// depths (MinCoCForBound, renderer.cpp:1602-1617: 0 across the focal plane,
// else the nearer end's least radius).
float adjusted_rate(const Node& n, const Ctx& c) {
    const Shape& s = *c.s;
    const rh_prim& p = c.prims[n.prim];
    float rate = p.shading_rate;
    if (s.dof) {
        const float z1 = n.rb.lo.z, z2 = n.rb.hi.z;
        const float focalDist = 1 / s.dofOneOverF;
        float minCoC = 0.0f;
        if (!((z1 - focalDist) * (z2 - focalDist) < 0)) {
            const float minBlur = mn(std::fabs(1 / z1 - s.dofOneOverF), std::fabs(1 / z2 - s.dofOneOverF));
            minCoC = s.dofMult * mn(s.dofScaleX, s.dofScaleY) * minBlur;
        }
        const float focusFactor = p.focus_factor, areaRatio = 0.025f;
        rate *= mx(1.0f, areaRatio * focusFactor * minCoC * minCoC);
    }
    return rate;
}

// CqQuadric::EstimateGridSize + Diceable (quadrics.cpp:259-346)
bool sphere_diceable(Node& n, const Ctx& c) {
    if (!n.diceableFlag) return false;
    const int E = 8;
    float su = 0, sv = 0;
    m44 M = m_mul(c.s->diceM, c.objToCam[n.prim]);
    float sT[E + 1], cT[E + 1], sP[E + 1], cP[E + 1];
    sincos_grid(deg2rad(n.thetaMin), deg2rad(n.thetaMax), E + 1, sT, cT);
    sincos_grid(n.phiMin, n.phiMax, E + 1, sP, cP);
    f3 prev = mk(0, 0, 0), up[E];
    for (int i = 0; i < E; ++i) up[i] = mk(0, 0, 0);
    for (int v = 0; v <= E; ++v)
        for (int u = 0; u <= E; ++u) {
            f3 p = m_point(M, scale(n.radius, mk(cT[u] * cP[v], sT[u] * cP[v], sP[v])));
            if (v >= 1 && u >= 1) {
                su = mx(su, len2(sub(p, prev)));
                sv = mx(sv, len2(sub(up[u - 1], prev)));
            }
            if (u >= 1) up[u - 1] = prev;
            prev = p;
        }
    su = std::sqrt(su); sv = std::sqrt(sv);
    float srt = std::sqrt(adjusted_rate(n, c));
    n.uDice = (int)ceil_l(E * su / srt);
    n.vDice = (int)ceil_l(E * sv / srt);
    unsigned long too = (unsigned long)n.uDice * n.vDice;
    n.splitDir = n.uDice > n.vDice ? DIR_U : DIR_V;
    if (too > 10000) return false;
    if (n.uDice > 16.0f || n.vDice > 16.0f) return false;
    return true;
}

// CqSurfacePatchBicubic::Diceable (patch.cpp:300-379)
bool bicubic_diceable(Node& n, const Ctx& c) {
    if (!n.diceableFlag) return false;
    f3 h[16];
    for (int i = 0; i < 16; ++i) h[i] = m_point(c.s->diceM, n.cp[i]);
    float ul = 0, vl = 0;
    for (int u = 0; u < 16; u += 4)
        for (int k = 0; k < 3; ++k) {
            float d = len2(sub(h[u + k + 1], h[u + k]));
            if (d > ul) ul = d;
        }
    for (int v = 0; v < 4; ++v)
        for (int k = 0; k < 3; ++k) {
            float d = len2(sub(h[v + 4 * k + 4], h[v + 4 * k]));
            if (d > vl) vl = d;
        }
    float sr = adjusted_rate(n, c);
    ul = std::sqrt(ul / sr); vl = std::sqrt(vl / sr);
    n.splitDir = ul > vl ? DIR_U : DIR_V;
    ul *= 3; vl *= 3;
    n.uDice = (int)mx<long>(round_l(ul), 1);
    n.vDice = (int)mx<long>(round_l(vl), 1);
    if (ul < FLT_EPSILON || vl < FLT_EPSILON) { n.discard = true; return false; }
    const float g2 = 256.0f;
    return !(n.uDice > g2 || n.vDice > g2 || (n.uDice * n.vDice) > g2);
}

// CqSurfacePatchBilinear::Diceable (patch.cpp:671-736)
bool bilinear_diceable(Node& n, const Ctx& c) {
    if (!n.diceableFlag) return false;
    f3 h[4];
    for (int i = 0; i < 4; ++i) h[i] = m_point(c.s->diceM, n.cp[i]);
    float a = len2(sub(h[1], h[0])), b = len2(sub(h[3], h[2]));
    float ul = a > b ? a : b;
    a = len2(sub(h[2], h[0])); b = len2(sub(h[3], h[1]));
    float vl = a > b ? a : b;
    float sr = adjusted_rate(n, c);
    ul = std::sqrt(ul / sr); vl = std::sqrt(vl / sr);
    n.splitDir = ul > vl ? DIR_U : DIR_V;
    ul = (float)mx<long>(round_l(ul), 1);
    vl = (float)mx<long>(round_l(vl), 1);
    n.uDice = (int)ul; n.vDice = (int)vl;
    if (ul < FLT_EPSILON || vl < FLT_EPSILON) { n.discard = true; return false; }
    const float g2 = 256.0f;
    return !(n.uDice > g2 || n.vDice > g2 || (n.uDice * n.vDice) > g2);
}

inline f3 mid(f3 a, f3 b) { return scale(0.5f, add(a, b)); }
inline f3 half(f3 a) { return scale(0.5f, a); }
inline f3 quarter_sum(f3 a, f3 b) { return scale(0.25f, add(a, b)); }

// bicubicPatchNatSubdiv (patch.cpp:87-129) on dehomogenised points
void split_bicubic(const f3* P, f3* A, f3* B, bool u) {
    for (int k = 0; k < 4; ++k) {
        int i0, s;
        if (u) { i0 = 4 * k; s = 1; } else { i0 = k; s = 4; }
        const f3 p0 = P[i0], p1 = P[i0 + s], p2 = P[i0 + 2 * s], p3 = P[i0 + 3 * s];
        A[i0] = p0;
        A[i0 + s] = mid(p0, p1);
        A[i0 + 2 * s] = add(half(A[i0 + s]), quarter_sum(p1, p2));
        B[i0 + 3 * s] = p3;
        B[i0 + 2 * s] = mid(p2, p3);
        B[i0 + s] = add(half(B[i0 + 2 * s]), quarter_sum(p1, p2));
        A[i0 + 3 * s] = mid(A[i0 + 2 * s], B[i0 + s]);
        B[i0] = A[i0 + 3 * s];
    }
}

// surfaceNaturalSubdivide (surface.cpp:301-327)
void split_bilinear(const f3* P, f3* A, f3* B, bool u) {
    for (int i = 0; i < 4; ++i) { A[i] = P[i]; B[i] = P[i]; }
    if (u) { A[1] = B[0] = mid(P[0], P[1]); A[3] = B[2] = mid(P[2], P[3]); }
    else { A[2] = B[0] = mid(P[0], P[2]); A[3] = B[1] = mid(P[1], P[3]); }
}

// CullSurface + bucket range (imagebuffer.cpp:268-345, 353-446)
void post_geometry(Node& n, const f3* cp1, const Ctx& c) {
    const Shape& s = *c.s;
    const rh_options& o = s.o;
    box3 B = (n.type == RH_PRIM_SPHERE) ? sphere_camera_bound(n, c.objToCam[n.prim])
                                        : patch_camera_bound(n, n.type == RH_PRIM_BICUBIC ? 16 : 4);
    if (n.motion) {
        // CqDeformingSurface::Bound (surface.h:629-640): the keys' bounds
        // encapsulated into an empty box
        Node k1 = n;
        for (int i = 0; i < 16; ++i) k1.cp[i] = cp1[i];
        box3 B1 = (n.type == RH_PRIM_SPHERE) ? sphere_camera_bound(n, c.objToCam1[n.prim])
                                             : patch_camera_bound(k1, n.type == RH_PRIM_BICUBIC ? 16 : 4);
        box3 u;
        u.lo = mk(FLT_MAX, FLT_MAX, FLT_MAX);
        u.hi = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
        for (const box3* b : {&B, &B1}) {
            u.hi = mk(mx(u.hi.x, b->hi.x), mx(u.hi.y, b->hi.y), mx(u.hi.z, b->hi.z));
            u.lo = mk(mn(u.lo.x, b->lo.x), mn(u.lo.y, b->lo.y), mn(u.lo.z, b->lo.z));
        }
        B = u;
    }
    if (B.lo.z >= o.clip_far || B.hi.z <= o.clip_near) { n.culled = true; return; }
    if (B.lo.z <= FLT_EPSILON) {
        n.diceableFlag = false;
        if (n.splitCount > o.eyesplits) n.culled = true;
        return;
    }
    float z0 = B.lo.z, z1 = B.hi.z;
    B = box_xform(B, s.camToRaster);
    if (s.dof) {
        // grown by the larger circle of confusion at its depths (imagebuffer.cpp:312-322)
        float c0x, c0y, c1x, c1y;
        circle_of_confusion(s, z0, c0x, c0y);
        circle_of_confusion(s, z1, c1x, c1y);
        const float cocX = mx(c0x, c1x), cocY = mx(c0y, c1y);
        B.lo.x = B.lo.x - cocX;
        B.lo.y = B.lo.y - cocY;
        B.hi.x = B.hi.x + cocX;
        B.hi.y = B.hi.y + cocY;
    }
    B.lo.x = B.lo.x - o.filter_xwidth / 2.0f;
    B.lo.y = B.lo.y - o.filter_ywidth / 2.0f;
    B.hi.x = B.hi.x + o.filter_xwidth / 2.0f;
    B.hi.y = B.hi.y + o.filter_ywidth / 2.0f;
    if (B.lo.x > s.cropX1 || B.lo.y > s.cropY1 || B.hi.x < s.cropX0 || B.hi.y < s.cropY0) { n.culled = true; return; }
    B.lo.z = z0; B.hi.z = z1;
    n.rb = B;
    n.hasBound = true;
    n.xminb = static_cast<int>(B.lo.x) / o.bucket_x;
    n.yminb = static_cast<int>(B.lo.y) / o.bucket_y;
    n.xmaxb = static_cast<int>(B.hi.x) / o.bucket_x;
    n.ymaxb = static_cast<int>(B.hi.y) / o.bucket_y;
}

// Expand the whole split tree of one primitive (everything that does not
// depend on the bucket being processed).
void build_tree(const Ctx& c, int prim, Tree& tr) {
    const rh_prim& p = c.prims[prim];
    std::vector<Node>& t = tr.n;
    Node root;
    root.type = p.type;
    root.prim = prim;
    root.motion = p.nkeys == 2;
    t.clear();
    tr.cp1.clear();
    if (p.type == RH_PRIM_SPHERE) {
        root.radius = p.sphere[0];
        root.thetaMin = 0.0f;
        root.thetaMax = p.sphere[3];
        root.phiMin = (float)std::asin((double)(p.sphere[1] / root.radius));
        root.phiMax = (float)std::asin((double)(p.sphere[2] / root.radius));
    } else {
        int k = p.type == RH_PRIM_BICUBIC ? 16 : 4;
        for (int i = 0; i < k; ++i) root.cp[i] = mk(p.P[i][0], p.P[i][1], p.P[i][2]);
    }
    if (root.motion) {
        tr.cp1.assign(16, mk(0, 0, 0));
        if (p.type != RH_PRIM_SPHERE)
            for (int i = 0; i < (p.type == RH_PRIM_BICUBIC ? 16 : 4); ++i) tr.cp1[i] = mk(p.P1[i][0], p.P1[i][1], p.P1[i][2]);
    }
    t.reserve(16);
    t.push_back(root);
    for (size_t i = 0; i < t.size(); ++i) {
        post_geometry(t[i], root.motion ? &tr.cp1[16 * i] : nullptr, c);
        if (t[i].culled) continue;
        Node& n = t[i];
        bool d;
        // A deforming surface asks key 0 (CqDeformingSurface::Diceable,
        // surface.h:698-707), whose own m_fDiceable is never cleared; a
        // discard key 0 decides stays with the key, so the deforming surface
        // splits instead.
        const bool outerFlag = n.diceableFlag, outerDiscard = n.discard;
        if (n.motion) n.diceableFlag = true;
        if (n.type == RH_PRIM_SPHERE) d = sphere_diceable(n, c);
        else if (n.type == RH_PRIM_BICUBIC) d = bicubic_diceable(n, c);
        else d = bilinear_diceable(n, c);
        if (n.motion) { n.diceableFlag = outerFlag; n.discard = outerDiscard; }
        if (d) { n.leaf = true; continue; }
        if (n.discard) continue;
        // CqSurface::Split (surface.cpp:529-577): two halves; an undiceable
        // (eye-split) surface splits each half once more.
        auto halves = [&](const Node& src, const f3* src1, Node& a, Node& b, f3* a1, f3* b1) {
            a = src; b = src;
            bool u = src.splitDir == DIR_U;
            if (src.type == RH_PRIM_SPHERE) {
                float pc = (float)((src.phiMin + src.phiMax) * 0.5);
                float tc = (float)((src.thetaMin + src.thetaMax) * 0.5);
                if (u) { a.thetaMax = tc; b.thetaMin = tc; } else { a.phiMax = pc; b.phiMin = pc; }
            } else if (src.type == RH_PRIM_BICUBIC) {
                split_bicubic(src.cp, a.cp, b.cp, u);
                if (src.motion) split_bicubic(src1, a1, b1, u);
            } else {
                split_bilinear(src.cp, a.cp, b.cp, u);
                if (src.motion) split_bilinear(src1, a1, b1, u);
            }
            for (Node* k : {&a, &b}) {
                k->splitDir = u ? DIR_V : DIR_U;
                k->splitCount = src.splitCount + 1;
                k->diceableFlag = true; k->discard = false; k->culled = false;
                k->hasBound = false; k->leaf = false; k->kid0 = -1; k->nkids = 0;
            }
        };
        Node kids[4];
        f3 kid1[4][16];
        int nk;
        const f3* own1 = root.motion ? &tr.cp1[16 * i] : nullptr;
        // (a deforming surface's keys split once: CqDeformingSurface::Split,
        // surface.h:657-692)
        if (!t[i].diceableFlag && t[i].type != RH_PRIM_BILINEAR && !t[i].motion) {
            Node a, b;
            f3 a1[16], b1[16];
            halves(t[i], own1, a, b, a1, b1);
            halves(a, a1, kids[0], kids[1], kid1[0], kid1[1]);
            halves(b, b1, kids[2], kids[3], kid1[2], kid1[3]);
            nk = 4;
        } else {
            halves(t[i], own1, kids[0], kids[1], kid1[0], kid1[1]);
            nk = 2;
        }
        t[i].kid0 = (int)t.size();
        t[i].nkids = nk;
        for (int q = 0; q < nk; ++q) {
            t.push_back(kids[q]);
            if (root.motion) tr.cp1.insert(tr.cp1.end(), kid1[q], kid1[q] + 16);
        }
    }
}

// The bucket loop's view of a node: what posting, popping, culling and
// reposting read, packed so that the loop walks a few megabytes.
enum { SN_LEAF = 1, SN_BOUND = 2, SN_CULLED = 4, SN_MOTION = 8, SN_CULLABLE = 16 };
struct SimNode {
    float z, subMin;            // cached bound's minimum z; least in the subtree
    float bx0, by0, bx1, by1;   // cached raster bound in x and y
    int16_t x0, y0, x1, y1;     // PostSurface's bucket range, clamped to the region
    uint32_t kid0;              // forest index of the first child
    int32_t prim;
    uint32_t param;             // a leaf's dicing parameters in Forest::params
    uint32_t pairStart;         // the node's first (surface, bucket) cull pair
    uint16_t cu, cv;            // dice sizes of a leaf
    uint8_t nkids, flags, type;
};

// (surface, bucket) pairs an occlusion test may ask about a node: every
// bucket of its range when it has a bound (PostSurface's range,
// imagebuffer.cpp:398-407, clamped to the region: posting and reposting keep
// it inside), none otherwise.
uint64_t node_pairs(const Shape& s, const Node& nd) {
    if (!nd.hasBound || nd.culled) return 0;
    const int x0 = clampv(nd.xminb, s.bx0, s.bx1 - 1), y0 = clampv(nd.yminb, s.by0, s.by1 - 1);
    const int x1 = clampv(nd.xmaxb, s.bx0, s.bx1 - 1), y1 = clampv(nd.ymaxb, s.by0, s.by1 - 1);
    return (x1 >= x0 && y1 >= y0) ? (uint64_t)(x1 - x0 + 1) * (uint64_t)(y1 - y0 + 1) : 0;
}

// Floats of a leaf's dicing parameters (layouts: front.h).
uint32_t leaf_param_floats(const Node& nd) {
    const uint32_t keys = nd.motion ? 2 : 1;
    const uint32_t pk = nd.type == RH_PRIM_SPHERE ? (uint32_t)(18 + 2 * (nd.uDice + 1) + 2 * (nd.vDice + 1))
                                                  : (nd.type == RH_PRIM_BICUBIC ? 48u : 12u);
    return pk * keys + (keys == 2 ? 2 : 0);
}

// A leaf's dicing parameters: each key in turn, then a deforming surface's
// two key times (its keys share the dice sizes, CopySplitInfo).
void write_leaf_params(const Ctx& c, const Tree& tr, size_t local, float* W) {
    const Node& nd = tr.n[local];
    for (int key = 0; key < (nd.motion ? 2 : 1); ++key) {
        if (nd.type == RH_PRIM_SPHERE) {
            const m44& M = key ? c.objToCam1[nd.prim] : c.objToCam[nd.prim];
            *W++ = nd.radius;
            for (int r = 0; r < 4; ++r)
                for (int cc = 0; cc < 4; ++cc) *W++ = M.e[r][cc];
            *W++ = M.id ? 1.0f : 0.0f;
            sincos_grid(deg2rad(nd.thetaMin), deg2rad(nd.thetaMax), nd.uDice + 1, W, W + nd.uDice + 1);
            W += 2 * (nd.uDice + 1);
            sincos_grid(nd.phiMin, nd.phiMax, nd.vDice + 1, W, W + nd.vDice + 1);
            W += 2 * (nd.vDice + 1);
        } else {
            const int kk = nd.type == RH_PRIM_BICUBIC ? 16 : 4;
            const f3* cp = key ? &tr.cp1[16 * local] : nd.cp;
            for (int j = 0; j < kk; ++j) { *W++ = cp[j].x; *W++ = cp[j].y; *W++ = cp[j].z; }
        }
    }
    if (nd.motion) {
        *W++ = c.prims[nd.prim].key_time[0];
        *W++ = c.prims[nd.prim].key_time[1];
    }
}

// A bucket heap entry carries the comparator's inputs so that heap moves
// and comparisons do not touch the nodes.
struct HeapEnt {
    float z;                    // cached raster bound's minimum z
    float subMin;               // least minimum z in the node's subtree
    bool hasBound;
    uint32_t nd;                // forest index
};

struct HeapCmp {  // closest_surface (bucket.h:164-176)
    bool operator()(const HeapEnt& a, const HeapEnt& b) const {
        if (a.hasBound && b.hasBound) return a.z > b.z;
        return true;
    }
};

}  // namespace

struct Forest {
    Ctx c;
    std::vector<Tree> trees;               // one split tree per primitive, parents first
    std::vector<uint64_t> nodeBase;        // primitives + 1: forest index of each tree's root
    std::vector<SimNode> sim;              // every node, in forest order
    std::unique_ptr<float[]> params;       // every leaf's dicing parameters, in forest order
    uint64_t nparams = 0;
    uint64_t leaves = 0;
    uint64_t npairs = 0;                   // cull pairs of all nodes (SimNode::pairStart)
    int threads = 1;
};

const float* forest_params(const Forest& f, uint64_t& n) {
    n = f.nparams;
    return f.params.get();
}

void ForestFree::operator()(Forest* f) const { delete f; }

bool build_forest(const Shape& s, const rh_prim* prims, int n, ForestPtr& out, std::string& err, int threads) {
    ForestPtr fp(new Forest());
    Forest& f = *fp;

    Ctx& c = f.c;
    c.s = &s;
    c.prims = prims;
    c.objToCam.assign(n, m_identity());
    c.objToCam1.assign(n, m_identity());
    for (int i = 0; i < n; ++i) {
        const rh_prim& p = prims[i];
        if (p.type != RH_PRIM_SPHERE && p.type != RH_PRIM_BICUBIC && p.type != RH_PRIM_BILINEAR) {
            err = "unknown surface type at index " + std::to_string(i);
            return false;
        }
        if (p.nkeys != 1 && p.nkeys != 2) { err = "nkeys must be 1 or 2"; return false; }
        if (p.nkeys == 2 && s.dof) {
            err = "depth of field together with motion keys is not supported";
            return false;
        }
        if (p.nkeys == 2 && !(p.key_time[0] < p.key_time[1])) {
            err = "motion key times must increase";
            return false;
        }
        if (!(p.shading_rate > 0)) { err = "shading rate must be positive"; return false; }
        if (p.matte != RH_MATTE_NONE && p.matte != RH_MATTE && p.matte != RH_MATTE_ALPHA) {
            err = "matte must be RH_MATTE_NONE, RH_MATTE or RH_MATTE_ALPHA";
            return false;
        }
        if (p.shading_interpolation != RH_SHADING_SMOOTH && p.shading_interpolation != RH_SHADING_CONSTANT) {
            err = "shading_interpolation must be RH_SHADING_SMOOTH or RH_SHADING_CONSTANT";
            return false;
        }
        if (p.type == RH_PRIM_SPHERE) {
            if (!(p.sphere[0] != 0)) { err = "sphere radius must be non-zero"; return false; }
            if (p.sphere[3] != 360.0f || p.sphere[1] != -p.sphere[0] || p.sphere[2] != p.sphere[0]) {
                // Partial spheres follow the same code; the CqSphere constructor's
                // zmin/zmax clamp is not restated, so keep them inside the radius.
                if (std::fabs(p.sphere[1]) > std::fabs(p.sphere[0]) || std::fabs(p.sphere[2]) > std::fabs(p.sphere[0])) {
                    err = "sphere zmin/zmax outside the radius";
                    return false;
                }
            }
            m44 M;
            bool ident = true;
            for (int r = 0; r < 4; ++r)
                for (int cc = 0; cc < 4; ++cc) {
                    M.e[r][cc] = p.xform[0][4 * r + cc];
                    if (M.e[r][cc] != (r == cc ? 1.0f : 0.0f)) ident = false;
                }
            M.id = ident;
            c.objToCam[i] = M;
            if (p.nkeys == 2) {
                m44 M1;
                bool id1 = true;
                for (int r = 0; r < 4; ++r)
                    for (int cc = 0; cc < 4; ++cc) {
                        M1.e[r][cc] = p.xform[1][4 * r + cc];
                        if (M1.e[r][cc] != (r == cc ? 1.0f : 0.0f)) id1 = false;
                    }
                M1.id = id1;
                c.objToCam1[i] = M1;
            }
        }
    }
    // Split trees, one per primitive, in parallel; each tree's leaf
    // parameter floats in paramBase[i + 1], its leaves in leafCount[i].
    std::vector<uint64_t> paramBase(n + 1, 0), pairBase(n + 1, 0);
    std::vector<uint32_t> leafCount(n, 0);
    f.trees.assign(n, Tree());
    f.threads = std::max(1, threads);
    auto work = [&](int i) {
        {
            build_tree(c, i, f.trees[i]);
            std::vector<Node>& t = f.trees[i].n;
            // the least minimum depth of each node's subtree (children follow
            // their parents in the tree)
            uint64_t pf = 0, pp = 0;
            uint32_t nl = 0;
            for (size_t j = t.size(); j-- > 0;) {
                Node& nd = t[j];
                float m = nd.hasBound ? nd.rb.lo.z : FLT_MAX;
                for (int q = 0; q < nd.nkids; ++q) m = mn(m, t[nd.kid0 + q].subMin);
                nd.subMin = m;
                if (nd.leaf) {
                    pf += leaf_param_floats(nd);
                    nl++;
                }
                pp += node_pairs(s, nd);
            }
            paramBase[i + 1] = pf;
            pairBase[i + 1] = pp;
            leafCount[i] = nl;
        }
    };
    run_parallel(n, threads, work);
    f.nodeBase.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) f.nodeBase[i + 1] = f.nodeBase[i] + f.trees[i].n.size();
    if (f.nodeBase[n] >= 0xffffffffull) {
        err = "the scene splits into more than 2^32 surfaces";
        return false;
    }
    // each tree's leaf parameters start after the previous trees'
    for (int i = 0; i < n; ++i) {
        paramBase[i + 1] += paramBase[i];
        pairBase[i + 1] += pairBase[i];
        f.leaves += leafCount[i];
    }
    f.npairs = pairBase[n];
    if (f.npairs >= 0xffffffffull) {
        err = "the scene's occlusion tests exceed 2^32 (surface, bucket) pairs";
        return false;
    }
    if (paramBase[n] >= 0xffffffffull) {
        err = "the scene's dicing parameters exceed 2^32 floats";
        return false;
    }
    // (not zeroed: every float is written below)
    f.nparams = paramBase[n];
    f.params.reset(new float[std::max<uint64_t>(f.nparams, 1)]);
    // the bucket loop's packed nodes
    f.sim.assign(f.nodeBase[n], SimNode());
    auto pack = [&](int i) {
        {
            const std::vector<Node>& t = f.trees[i].n;
            uint64_t po = paramBase[i], pq = pairBase[i];
            for (size_t j = 0; j < t.size(); ++j) {
                const Node& nd = t[j];
                SimNode& q = f.sim[f.nodeBase[i] + j];
                q.type = (uint8_t)nd.type;
                q.param = (uint32_t)po;
                q.pairStart = (uint32_t)pq;
                pq += node_pairs(s, nd);
                if (nd.leaf) {
                    write_leaf_params(c, f.trees[i], j, &f.params[po]);
                    po += leaf_param_floats(nd);
                }
                q.z = nd.rb.lo.z;
                q.subMin = nd.subMin;
                q.bx0 = nd.rb.lo.x; q.by0 = nd.rb.lo.y; q.bx1 = nd.rb.hi.x; q.by1 = nd.rb.hi.y;
                int x0 = 0, y0 = 0, x1 = 0, y1 = 0;
                if (nd.diceableFlag) { x0 = nd.xminb; y0 = nd.yminb; x1 = nd.xmaxb; y1 = nd.ymaxb; }
                q.x0 = (int16_t)clampv(x0, s.bx0, s.bx1 - 1); q.y0 = (int16_t)clampv(y0, s.by0, s.by1 - 1);
                q.x1 = (int16_t)clampv(x1, s.bx0, s.bx1 - 1); q.y1 = (int16_t)clampv(y1, s.by0, s.by1 - 1);
                q.kid0 = nd.nkids ? (uint32_t)(f.nodeBase[i] + nd.kid0) : 0u;
                q.prim = nd.prim;
                q.cu = (uint16_t)nd.uDice; q.cv = (uint16_t)nd.vDice;
                q.nkids = (uint8_t)nd.nkids;
                q.flags = (uint8_t)((nd.leaf ? SN_LEAF : 0) | (nd.hasBound ? SN_BOUND : 0) |
                                    (nd.culled ? SN_CULLED : 0) |
                                    (prims[i].no_cull_hidden ? 0 : SN_CULLABLE) | (nd.motion ? SN_MOTION : 0));
            }
        }
    };
    run_parallel(n, threads, pack);
    // the caller's shape and surfaces are passed again to simulate
    c.s = nullptr;
    c.prims = nullptr;
    out = std::move(fp);
    return true;
}

template <typename Fn> static void parallel_for(int n, int threads, Fn fn);

void cull_pairs(const Shape& s, const Forest& f, CullPairs& out) {
    (void)s;
    const size_t nn = f.sim.size();
    out.bound.resize(nn * 4);
    out.range.resize(nn * 4);
    out.start.resize(nn + 1);
    const int nchunk = (int)std::min<size_t>((size_t)std::max(1, f.threads), nn / 4096 + 1);
    const size_t per = (nn + nchunk - 1) / nchunk;
    parallel_for(nchunk, f.threads, [&](int ch) {
        for (size_t g = ch * per; g < std::min(nn, (ch + 1) * per); ++g) {
            const SimNode& q = f.sim[g];
            const bool tested = (q.flags & SN_BOUND) && !(q.flags & SN_CULLED);
            out.bound[4 * g] = tested ? q.bx0 : 0.0f; out.bound[4 * g + 1] = tested ? q.by0 : 0.0f;
            out.bound[4 * g + 2] = tested ? q.bx1 : 0.0f; out.bound[4 * g + 3] = tested ? q.by1 : 0.0f;
            out.range[4 * g] = tested ? q.x0 : 0; out.range[4 * g + 1] = tested ? q.y0 : 0;
            out.range[4 * g + 2] = tested ? q.x1 : -1; out.range[4 * g + 3] = tested ? q.y1 : -1;
            out.start[g] = q.pairStart;
        }
    });
    out.start[nn] = f.npairs;
}

uint64_t cull_trees(const Shape& s, std::vector<CullTreeDev>& out) {
    out.assign(s.buckets.size(), CullTreeDev());
    uint64_t pyr = 0;
    for (size_t k = 0; k < s.buckets.size(); ++k) {
        const auto& b = s.buckets[k];
        CullTreeDev& t = out[k];
        // CqOcclusionTree::setupTree (occlusion.cpp:54-107)
        const int w = b.sx1 - b.sx0, h = b.sy1 - b.sy0;
        const int numX = w * s.nx, numY = h * s.ny;
        int depthX = 0, depthY = 0;
        if (numX > 0 && numY > 0) {
            // aqsis's log2: 1.4426950408889633 * std::log(float) (math.h:156-160)
            depthX = (int)ceil_l((float)(1.4426950408889633 * (double)std::log((float)numX)));
            depthY = (int)ceil_l((float)(1.4426950408889633 * (double)std::log((float)numY)));
            if (depthX < depthY) depthX = depthY - 1;
            else if (depthY < depthX) depthY = depthX - 1;
        }
        t.dx = depthX; t.dy = depthY;
        t.cw = numX; t.ch = numY;
        t.sx0 = b.sx0; t.sy0 = b.sy0;
        t.x0 = (float)b.sx0; t.y0 = (float)b.sy0;
        if (numX > 0 && numY > 0) {
            const float scx = float(1 << depthX) / numX, scy = float(1 << depthY) / numY;
            t.x1 = t.x0 + (float)w * scx;
            t.y1 = t.y0 + (float)h * scy;
        } else {
            t.x1 = t.x0; t.y1 = t.y0;
        }
        t.pyr = (uint32_t)pyr;
        t.levels = 0;
        if (numX > 0 && numY > 0) {
            for (int l = 0;; ++l) {
                const int lw = cull_level_dim(numX, l), lh = cull_level_dim(numY, l);
                pyr += (uint64_t)lw * lh;
                t.levels = l + 1;
                if (lw == 1 && lh == 1) break;
            }
        }
    }
    return pyr;
}

// Run fn(i) for i in [0, n) on up to `threads` threads.
template <typename Fn> static void parallel_for(int n, int threads, Fn fn) {
    run_parallel(n, threads, std::function<void(int)>(fn));
}

static void finish_schedule(const Shape& s, const Forest& f, Schedule& out);

// The grid of a leaf diced in region bucket k, its first micropolygon id
// and vertex given; its parameters are the forest's.
static bool leaf_ok(const SimNode& q, std::string& err) {
    if ((q.cu + 1) * (q.cv + 1) > 514 || q.cu < 1 || q.cv < 1) {
        err = "internal: grid size outside the dicing limits";
        return false;
    }
    return true;
}
static uint32_t leaf_vertices(const SimNode& q) {
    return (uint32_t)((q.cu + 1) * (q.cv + 1)) * ((q.flags & SN_MOTION) ? 2u : 1u);
}
static GridRec make_grid(const Shape& s, const SimNode& q, int k, uint32_t mpBase, uint32_t vtxBase) {
    const rh_options& o = s.o;
    auto bidx = [](float v, int bs) { return (int)std::floor((double)v / bs); };
    const bool motion = (q.flags & SN_MOTION) != 0;
    GridRec gr;
    gr.type = q.type | (motion ? GRID_MOTION : 0);
    gr.cu = q.cu; gr.cv = q.cv; gr.prim = q.prim;
    gr.mpBase = mpBase;
    gr.vtxBase = vtxBase;
    gr.paramOff = q.param;
    gr.diceBucket = k;
    // Candidate buckets: the cached raster bound (already grown by half the
    // filter width) plus a one-pixel guard band.
    gr.cbx0 = (int16_t)clampv(bidx(q.bx0 - 1.0f, o.bucket_x), s.bx0, s.bx1 - 1);
    gr.cby0 = (int16_t)clampv(bidx(q.by0 - 1.0f, o.bucket_y), s.by0, s.by1 - 1);
    gr.cbx1 = (int16_t)clampv(bidx(q.bx1 + 1.0f, o.bucket_x), s.bx0, s.bx1 - 1);
    gr.cby1 = (int16_t)clampv(bidx(q.by1 + 1.0f, o.bucket_y), s.by0, s.by1 - 1);
    gr.px0 = q.bx0 - 2.0f; gr.py0 = q.by0 - 2.0f;
    gr.px1 = q.bx1 + 2.0f; gr.py1 = q.by1 + 2.0f;
    return gr;
}

// Record a leaf diced in region bucket k after the grids diced so far.
static bool grid_record(const Shape& s, const SimNode& q, int k, uint32_t& mpCount, uint32_t& vtxCount,
                        Schedule& out, std::string& err) {
    if (!leaf_ok(q, err)) return false;
    out.grids.push_back(make_grid(s, q, k, mpCount, vtxCount));
    if (q.flags & SN_MOTION) out.motion = true;
    mpCount += (uint32_t)(q.cu * q.cv);
    vtxCount += leaf_vertices(q);
    return true;
}

bool simulate(const Shape& s, const rh_prim* prims, int n, const Forest& f, const CullInput* cull,
              const CullPairs* pairs, Schedule& out, std::string& err) {
    (void)pairs;   // the pairs' ranges and starts are the packed nodes'
    const std::vector<SimNode>& sn = f.sim;
    std::vector<CullTreeDev> ctrees;
    if (cull) cull_trees(s, ctrees);
    // Bucket loop: post roots in order, pop nearest first, cull (against
    // the bucket's occlusion tree) or dice or split, post children.
    const int nb = (int)s.buckets.size();
    std::vector<std::vector<HeapEnt>> heaps(nb);
    std::vector<uint8_t> processed(nb, 0);
    out = Schedule();
    auto push = [&](int k, const HeapEnt& ent) {
        heaps[k].push_back(ent);
        std::push_heap(heaps[k].begin(), heaps[k].end(), HeapCmp());
    };
    auto post = [&](uint32_t g) {
        out.surfaces++;
        const SimNode& q = sn[g];
        if (q.flags & SN_CULLED) return;
        const int x0 = q.x0, y0 = q.y0, x1 = q.x1, y1 = q.y1;
        int k = s.bucketIndex(x0, y0);
        const HeapEnt ent{q.z, q.subMin, (q.flags & SN_BOUND) != 0, g};
        if (!processed[k]) {
            push(k, ent);
            return;
        }
        int yb = y0, xb = x0 + 1;
        while (yb <= y1) {
            while (xb <= x1) {
                int kk = s.bucketIndex(xb, yb);
                if (!processed[kk]) {
                    push(kk, ent);
                    return;
                }
                ++xb;
            }
            xb = x0;
            ++yb;
        }
    };
    // CqImageBuffer::RepostSurface (imagebuffer.cpp:449-483): the bucket to
    // the right when the bound reaches it, else the next row's bucket under
    // the bound's left side; dropped past the last row.
    auto repost = [&](int k, uint32_t g) {
        const SimNode& q = sn[g];
        const auto& b = s.buckets[k];
        const int col = s.bx0 + k % s.nbr, row = s.by0 + k / s.nbr;
        int nextX = col + 1, nextY = row;
        const HeapEnt ent{q.z, q.subMin, (q.flags & SN_BOUND) != 0, g};
        if (nextX < s.bx1 && q.bx1 >= (float)(b.x + b.w)) {
            push(s.bucketIndex(nextX, nextY), ent);
            return;
        }
        ++nextY;
        nextX = (int)mx<long>((long)s.bx0, floor_l(q.bx0) / s.o.bucket_x);
        if (nextX < s.bx1 && nextY < s.by1 && q.by1 >= (float)(b.y + b.h)) push(s.bucketIndex(nextX, nextY), ent);
    };
    for (int i = 0; i < n; ++i) post((uint32_t)f.nodeBase[i]);
    // Each grid is recorded as it is diced: its ids and vertices follow the
    // dicing order, its parameters are the forest's.
    out.grids.reserve(f.leaves);
    uint32_t mpCount = 0, vtxCount = 0;
    for (int k = 0; k < nb; ++k) {
        auto& h = heaps[k];
        while (!h.empty()) {
            const uint32_t g = h.front().nd;
            std::pop_heap(h.begin(), h.end(), HeapCmp());
            h.pop_back();
            // the next surface's node, while this one is handled
            if (!h.empty()) __builtin_prefetch(&sn[h.front().nd]);
            const SimNode& q = sn[g];
            // RenderSurface's occlusion cull (bucketprocessor.cpp:945-957):
            // Attribute "cull" "hidden" (default 1) and a cached bound.
            if (cull && (q.flags & (SN_BOUND | SN_CULLABLE)) == (SN_BOUND | SN_CULLABLE)) {
                const int col = s.bx0 + k % s.nbr, row = s.by0 + k / s.nbr;
                if (col < q.x0 || col > q.x1 || row < q.y0 || row > q.y1) {
                    err = "internal: occlusion test outside a surface's bucket range";
                    return false;
                }
                const uint64_t pi = (uint64_t)q.pairStart + (uint64_t)(row - q.y0) * (uint64_t)(q.x1 - q.x0 + 1) +
                                    (uint64_t)(col - q.x0);
                const float z = q.z, M = cull->pairMax[pi];
                bool culled = M < z;
                auto fx = std::lower_bound(cull->fixed.begin(), cull->fixed.end(), std::make_pair(pi, (uint8_t)0));
                if (fx != cull->fixed.end() && fx->first == pi) {
                    culled = fx->second != 0;
                } else {
                    // The final depths settle the test unless they lie within
                    // the margin of z, or of the least depth a surface still in
                    // the heap (or a descendant) can reach; a heap holding a
                    // surface without a bound has no depth order at all.
                    bool settled;
                    if (culled) {
                        const float margin = kCullMargin * std::fabs(z);
                        float lim = z;
                        bool unordered = false;
                        for (const HeapEnt& e : h) {
                            if (!e.hasBound) unordered = true;
                            else lim = mn(lim, e.subMin);
                        }
                        settled = !unordered && M < lim - margin;
                    } else {
                        settled = M >= z + kKeepMargin * std::fabs(z);
                    }
                    if (!settled) {
                        CullProbe pr;
                        pr.pair = pi;
                        pr.bucket = k;
                        pr.mpCut = mpCount;
                        pr.z = z;
                        pr.cx0 = pr.cx1 = pr.cy0 = pr.cy1 = 0;
                        pr.kind = ctrees[k].levels ? cull_cells(ctrees[k], q.bx0, q.by0, q.bx1, q.by1, pr.cx0,
                                                                pr.cx1, pr.cy0, pr.cy1)
                                                   : 0;
                        pr.culled = culled ? 1 : 0;
                        out.probes.push_back(pr);
                    }
                }
                if (culled) {
                    out.culled++;
                    repost(k, g);
                    continue;
                }
            }
            if (q.flags & SN_LEAF) {
                if (!grid_record(s, q, k, mpCount, vtxCount, out, err)) return false;
            } else {
                // children live in the same tree as their parent
                for (int j = 0; j < q.nkids; ++j) post(q.kid0 + (uint32_t)j);
            }
        }
        processed[k] = 1;
    }
    out.micropolygons = mpCount;
    out.vertices = vtxCount;
    out.paramFloats = f.nparams;
    finish_schedule(s, f, out);
    return true;
}

// The bucket lists of a schedule's grids (in dicing order).
static void finish_schedule(const Shape& s, const Forest& f, Schedule& out) {
    const int nb = (int)s.buckets.size();
    // Candidate lists per bucket (grids in dicing order), only buckets not
    // yet processed when the grid was diced: counted per chunk of grids in
    // parallel, then each chunk writes its part of every bucket's list.
    const uint32_t ng = (uint32_t)out.grids.size();
    const int nchunk = (int)std::min<uint32_t>((uint32_t)std::max(1, f.threads), ng / 1024 + 1);
    const uint32_t per = (ng + nchunk - 1) / nchunk;
    std::vector<uint32_t> cnt((size_t)nchunk * nb, 0);
    auto each_cand = [&](uint32_t gi, auto fn) {
        const auto& g = out.grids[gi];
        for (int by = g.cby0; by <= g.cby1; ++by)
            for (int bx = g.cbx0; bx <= g.cbx1; ++bx) {
                const int kk = s.bucketIndex(bx, by);
                if (kk >= g.diceBucket) fn(kk);
            }
    };
    parallel_for(nchunk, f.threads, [&](int ch) {
        uint32_t* C = &cnt[(size_t)ch * nb];
        for (uint32_t gi = ch * per; gi < std::min(ng, (ch + 1) * per); ++gi) each_cand(gi, [&](int kk) { C[kk]++; });
    });
    out.candStart.assign(nb + 1, 0);
    {
        uint32_t run = 0;
        for (int k = 0; k < nb; ++k) {
            out.candStart[k] = run;
            for (int ch = 0; ch < nchunk; ++ch) {
                const uint32_t c0 = cnt[(size_t)ch * nb + k];
                cnt[(size_t)ch * nb + k] = run;      // the chunk's first slot in bucket k
                run += c0;
            }
        }
        out.candStart[nb] = run;
    }
    out.candGrid.resize(out.candStart[nb]);
    parallel_for(nchunk, f.threads, [&](int ch) {
        uint32_t* C = &cnt[(size_t)ch * nb];
        for (uint32_t gi = ch * per; gi < std::min(ng, (ch + 1) * per); ++gi)
            each_cand(gi, [&](int kk) { out.candGrid[C[kk]++] = gi; });
    });
    // Each task's list capacity, bucket by bucket in parallel: a
    // micropolygon enters a task's list only if its pixel window meets the
    // task, and a static one's window lies inside the grid's guarded bound.
    // A moving one's window (its sub-bounds) may reach past its keys' bound
    // to anywhere in the bucket.
    const int ntask = (int)s.tasks.size();
    out.capStart.assign(ntask + 1, 0);
    out.segBucket.resize(out.candGrid.size());
    parallel_for(nb, f.threads, [&](int k) {
        for (uint32_t e = out.candStart[k]; e < out.candStart[k + 1]; ++e) {
            out.segBucket[e] = (uint32_t)k;
            const auto& g = out.grids[out.candGrid[e]];
            const bool anyTask = (g.type & GRID_MOTION) != 0;
            for (int t = s.taskStart[k]; t < s.taskStart[k + 1]; ++t) {
                const auto& T = s.tasks[t];
                if (anyTask || (g.px0 < (float)T.x1 && g.px1 > (float)T.x0 && g.py0 < (float)T.y1 &&
                                g.py1 > (float)T.y0))
                    out.capStart[t + 1] += (uint64_t)(g.cu * g.cv);
            }
        }
    });
    for (int t = 0; t < ntask; ++t) out.capStart[t + 1] += out.capStart[t];
}

bool depth_schedule(const Shape& s, const rh_prim* prims, int n, const Forest& f, Schedule& out, std::string& err) {
    (void)prims; (void)n;
    const std::vector<SimNode>& sn = f.sim;
    const int nb = (int)s.buckets.size();
    out = Schedule();
    out.surfaces = sn.size();
    // Every leaf, grouped by its bucket range's first bucket (where the
    // culling-free bucket loop dices it), in forest order within a bucket:
    // counted per chunk of nodes, then placed.
    const uint32_t nn = (uint32_t)sn.size();
    const int nchunk = (int)std::min<uint32_t>((uint32_t)std::max(1, f.threads), nn / 4096 + 1);
    const uint32_t per = (nn + nchunk - 1) / nchunk;
    std::vector<uint32_t> cnt((size_t)nchunk * nb, 0);
    auto leafBucket = [&](const SimNode& q) { return s.bucketIndex(q.x0, q.y0); };
    parallel_for(nchunk, f.threads, [&](int ch) {
        uint32_t* C = &cnt[(size_t)ch * nb];
        for (uint32_t g = ch * per; g < std::min(nn, (ch + 1) * per); ++g)
            if (sn[g].flags & SN_LEAF) C[leafBucket(sn[g])]++;
    });
    uint32_t run = 0;
    for (int k = 0; k < nb; ++k)
        for (int ch = 0; ch < nchunk; ++ch) {
            const uint32_t c0 = cnt[(size_t)ch * nb + k];
            cnt[(size_t)ch * nb + k] = run;
            run += c0;
        }
    std::vector<uint32_t> order(run);
    parallel_for(nchunk, f.threads, [&](int ch) {
        uint32_t* C = &cnt[(size_t)ch * nb];
        for (uint32_t g = ch * per; g < std::min(nn, (ch + 1) * per); ++g)
            if (sn[g].flags & SN_LEAF) order[C[leafBucket(sn[g])]++] = g;
    });
    // ids and vertices: running sums over chunks of the order, then the records
    const int nrc = (int)std::min<uint32_t>((uint32_t)std::max(1, f.threads), run / 2048 + 1);
    const uint32_t rper = (run + nrc - 1) / nrc;
    std::vector<uint64_t> mpSum(nrc + 1, 0), vxSum(nrc + 1, 0);
    std::vector<uint8_t> bad(nrc, 0), moving(nrc, 0);
    parallel_for(nrc, f.threads, [&](int ch) {
        for (uint32_t i = ch * rper; i < std::min(run, (ch + 1) * rper); ++i) {
            const SimNode& q = sn[order[i]];
            if ((q.cu + 1) * (q.cv + 1) > 514 || q.cu < 1 || q.cv < 1) bad[ch] = 1;
            if (q.flags & SN_MOTION) moving[ch] = 1;
            mpSum[ch + 1] += (uint64_t)(q.cu * q.cv);
            vxSum[ch + 1] += leaf_vertices(q);
        }
    });
    for (int ch = 0; ch < nrc; ++ch) {
        if (bad[ch]) {
            err = "internal: grid size outside the dicing limits";
            return false;
        }
        if (moving[ch]) out.motion = true;
        mpSum[ch + 1] += mpSum[ch];
        vxSum[ch + 1] += vxSum[ch];
    }
    if (mpSum[nrc] >= 0xffffffffull || vxSum[nrc] >= 0xffffffffull) {
        err = "the schedule's micropolygons or vertices exceed 2^32";
        return false;
    }
    out.grids.resize(run);
    parallel_for(nrc, f.threads, [&](int ch) {
        uint32_t mp = (uint32_t)mpSum[ch], vx = (uint32_t)vxSum[ch];
        for (uint32_t i = ch * rper; i < std::min(run, (ch + 1) * rper); ++i) {
            const SimNode& q = sn[order[i]];
            out.grids[i] = make_grid(s, q, leafBucket(q), mp, vx);
            mp += (uint32_t)(q.cu * q.cv);
            vx += leaf_vertices(q);
        }
    });
    out.micropolygons = mpSum[nrc];
    out.vertices = vxSum[nrc];
    out.paramFloats = f.nparams;
    finish_schedule(s, f, out);
    return true;
}

bool make_schedule(const Shape& s, const rh_prim* prims, int n, Schedule& out, std::string& err, int threads) {
    ForestPtr f;
    if (!build_forest(s, prims, n, f, err, threads)) return false;
    return simulate(s, prims, n, *f, nullptr, nullptr, out, err);
}

}  // namespace rh
