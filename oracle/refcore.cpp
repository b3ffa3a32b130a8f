// ORACLE -- test infrastructure only (see refcore.h).
#include "refcore.h"

#include <utility>

namespace orc {

// init_genrand: libs/math/random.cpp:102-116.
void Rng::reseed(uint32_t s) {
    mt[0] = s;
    for (mti = 1; mti < 624; ++mti) {
        uint64_t prev = mt[mti - 1];
        mt[mti] = (uint32_t)((1812433253ULL * (prev ^ (prev >> 30)) + (uint64_t)mti) & 0xffffffffULL);
    }
}

// genrand_int32: libs/math/random.cpp:119-157.
uint32_t Rng::next() {
    static const uint32_t mag01[2] = {0x0u, 0x9908b0dfu};
    uint32_t y;
    if (mti >= 624) {
        if (mti == 625) reseed(5489u);
        int kk;
        for (kk = 0; kk < 624 - 397; ++kk) {
            y = (mt[kk] & 0x80000000u) | (mt[kk + 1] & 0x7fffffffu);
            mt[kk] = mt[kk + 397] ^ (y >> 1) ^ mag01[y & 0x1u];
        }
        for (; kk < 623; ++kk) {
            y = (mt[kk] & 0x80000000u) | (mt[kk + 1] & 0x7fffffffu);
            mt[kk] = mt[kk + (397 - 624)] ^ (y >> 1) ^ mag01[y & 0x1u];
        }
        y = (mt[623] & 0x80000000u) | (mt[0] & 0x7fffffffu);
        mt[623] = mt[396] ^ (y >> 1) ^ mag01[y & 0x1u];
        mti = 0;
    }
    y = mt[mti++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

// multiJitterIndices: libs/core/multijitter.cpp:83-128.
static void multi_jitter_indices(Rng& rng, int* idx, int numX, int numY) {
    for (int iy = 0; iy < numY; ++iy)
        for (int ix = 0; ix < numX; ++ix) {
            int which = 2 * (iy * numX + ix);
            idx[which] = iy;
            idx[which + 1] = ix;
        }
    for (int iy = 0; iy < numY; ++iy) {
        int ix = numX;
        while (ix > 1) {
            int ix2 = (int)rng.rint((uint32_t)ix);
            --ix;
            std::swap(idx[2 * (iy * numX + ix) + 1], idx[2 * (iy * numX + ix2) + 1]);
        }
    }
    for (int ix = 0; ix < numX; ++ix) {
        int iy = numY;
        while (iy > 1) {
            int iy2 = (int)rng.rint((uint32_t)iy);
            --iy;
            std::swap(idx[2 * (iy * numX + ix)], idx[2 * (iy2 * numX + ix)]);
        }
    }
}

// setupJitterPattern: libs/core/multijitter.cpp:130-202.
static void setup_jitter_pattern(Rng& rng, SamplerTables& t, int offset) {
    const int n = t.n;
    if (t.nx == 1 && t.ny == 1) {
        // CqVector2D(random.RandomFloat(), random.RandomFloat()): g++ on
        // x86-64 evaluates these constructor arguments right to left.
        float ry = rng.rfloat();
        float rx = rng.rfloat();
        t.pos2d[offset] = v2(rx, ry);
        t.pos1d[offset] = rng.rfloat();
    } else {
        std::vector<int> idx(2 * n);
        multi_jitter_indices(rng, idx.data(), t.nx, t.ny);
        float subPixelHeight = 1.0f / t.ny;
        float subPixelWidth = 1.0f / t.nx;
        float subcellWidth = 1.0f / n;
        int which = 0;
        for (int iy = 0; iy < t.ny; ++iy)
            for (int ix = 0; ix < t.nx; ++ix) {
                int xindex = idx[2 * which];
                int yindex = idx[2 * which + 1];
                // Two RandomFloat() calls inside one CqVector2D constructor
                // call (multijitter.cpp:166-168); g++ evaluates the y argument
                // first.
                float ry = rng.rfloat();
                float rx = rng.rfloat();
                float y = (yindex + ry) * subcellWidth + iy * subPixelHeight;
                float x = (xindex + rx) * subcellWidth + ix * subPixelWidth;
                t.pos2d[offset + which] = v2(x, y);
                ++which;
            }
    }
    float sample1d = 0;
    float delta1d = 1.0f / n;
    float random1d = rng.rfloat(delta1d);
    for (int i = 0; i < n; ++i) {
        float tt = sample1d + random1d;
        t.pos1d[offset + i] = tt;
        sample1d += delta1d;
    }
    for (int i = 0; i < n; ++i) t.shuffled[offset + i] = i;
    int j = n;
    while (j > 1) {
        int j2 = (int)rng.rint((uint32_t)j);
        --j;
        std::swap(t.shuffled[offset + j], t.shuffled[offset + j2]);
    }
}

// CqMultiJitteredSampler constructor: multijitter.h:90-101.
void build_multijitter(Rng& rng, int nx, int ny, SamplerTables& t) {
    t.nx = nx; t.ny = ny; t.n = nx * ny; t.cache = 250;
    t.pos2d.assign((size_t)t.n * t.cache, v2(0, 0));
    t.pos1d.assign((size_t)t.n * t.cache, 0.0f);
    t.shuffled.assign((size_t)t.n * t.cache, 0);
    for (int i = 0; i < t.cache; ++i) setup_jitter_pattern(rng, t, i * t.n);
    rng.reseed(19);
}

// CqGridSampler::setupGridPattern: libs/core/grid.cpp:37-63.
void build_grid_sampler(int nx, int ny, SamplerTables& t) {
    t.nx = nx; t.ny = ny; t.n = nx * ny; t.cache = 1;
    t.pos2d.assign(t.n, v2(0, 0));
    t.pos1d.assign(t.n, 0.0f);
    t.shuffled.assign(t.n, 0);
    float xScale = (float)(1.0 / nx);
    float yScale = (float)(1.0 / ny);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
            t.pos2d[j * nx + i] = v2((float)(xScale * (i + 0.5)), (float)(yScale * (j + 0.5)));
    // dt = 1/nSamples is integer division (grid.cpp:52)
    float dt = (float)(1 / t.n);
    float s = (float)(dt * 0.5);
    for (int i = 0; i < t.n; ++i) { t.pos1d[i] = s; s += dt; }
    for (int i = 0; i < t.n; ++i) t.shuffled[i] = i;
}

// CqMatrix::Rotate on an identity matrix: matrix.cpp:96-135.
Mat mat_rotate(float angle, V3 axis) {
    Mat R = mat_identity();
    if (angle != 0.0f && mag(axis) != 0.0f) {
        R.ident = false;
        V3 a = unit(axis);
        float s = (float)std::sin((double)angle);
        float c = (float)std::cos((double)angle);
        float t = 1.0f - c;
        R.m[0][0] = t * a.x * a.x + c;
        R.m[1][1] = t * a.y * a.y + c;
        R.m[2][2] = t * a.z * a.z + c;
        float txy = t * a.x * a.y, sz = s * a.z;
        R.m[0][1] = txy + sz;
        R.m[1][0] = txy - sz;
        float txz = t * a.x * a.z, sy = s * a.y;
        R.m[0][2] = txz - sy;
        R.m[2][0] = txz + sy;
        float tyz = t * a.y * a.z, sx = s * a.x;
        R.m[1][2] = tyz + sx;
        R.m[2][1] = tyz - sx;
    }
    return R;
}

// filters.cpp:72-342.  Expressions mix float arguments with double literals
// and double libm calls exactly as written there.
float filter_eval(int filter, float x, float y, float xwidth, float ywidth) {
    const double PI = 3.14159265358979323846;
    switch (filter) {
        case 0: {  // RiBoxFilter :129-141
            double a = (std::fabs(x) <= xwidth / 2.0) ? 1.0 : 0.0;
            double b = (std::fabs(y) <= ywidth / 2.0) ? 1.0 : 0.0;
            return (float)tmin(a, b);
        }
        case 1: {  // RiTriangleFilter :149-161
            float hxw = (float)(xwidth / 2.0);
            float hyw = (float)(ywidth / 2.0);
            float absx = std::fabs(x), absy = std::fabs(y);
            double a = absx <= hxw ? (double)((hxw - absx) / hxw) : 0.0;
            double b = absy <= hyw ? (double)((hyw - absy) / hyw) : 0.0;
            return (float)tmin(a, b);
        }
        case 2: {  // RiCatmullRomFilter :169-225 (RI spec 3.2 branch)
            float r2 = (x * x + y * y);
            float r = std::sqrt(r2);
            double v = (r >= 2.0) ? 0.0
                       : (r < 1.0) ? (3.0 * r * r2 - 5.0 * r2 + 2.0)
                                   : (-r * r2 + 5.0 * r2 - 8.0 * r + 4.0);
            return (float)v;
        }
        case 3: {  // RiGaussianFilter :72-109
            x /= xwidth;
            y /= ywidth;
            return (float)std::exp(-8.0 * (double)(x * x + y * y));
        }
        case 4: {  // RiSincFilter :233-278
            if (x != 0.0) {
                x = (float)(x * PI);
                x = (float)(std::cos(0.5 * x / xwidth) * std::sin((double)x) / x);
            } else x = 1.0f;
            if (y != 0.0) {
                y = (float)(y * PI);
                y = (float)(std::cos(0.5 * y / ywidth) * std::sin((double)y) / y);
            } else y = 1.0f;
            return x * y;
        }
        case 5: {  // RiDiskFilter :286-304
            double xx = x * x, yy = y * y;
            xwidth *= 0.5; ywidth *= 0.5;
            double d = xx / (xwidth * xwidth) + yy / (ywidth * ywidth);
            return d < 1.0 ? 1.0f : 0.0f;
        }
        case 6: {  // RiBesselFilter :312-342
            double xx = x * x, yy = y * y;
            xwidth *= 0.5; ywidth *= 0.5;
            double w = xx / (xwidth * xwidth) + yy / (ywidth * ywidth);
            if (w < 1.0) {
                double d = std::sqrt(xx + yy);
                if (d != 0.0) {
                    w = std::cos(0.5 * PI * std::sqrt(w));
                    return (float)(w * 2 * j1(PI * d) / d);
                }
                return (float)PI;
            }
            return 0.0f;
        }
        case 7: {  // RiMitchellFilter :117-122 via CqMitchellFilter :36-58
            const float B = 1 / 3.0f, C = 1 / 3.0f;
            float invX = 1.0f / xwidth, invY = 1.0f / ywidth;
            auto ev = [&](float t) -> float {
                t = std::fabs(2.f * t);
                if (t > 1.f)
                    return ((-B - 6 * C) * t * t * t + (6 * B + 30 * C) * t * t +
                            (-12 * B - 48 * C) * t + (8 * B + 24 * C)) * (1.f / 6.f);
                return ((12 - 9 * B - 6 * C) * t * t * t + (-18 + 12 * B + 6 * C) * t * t +
                        (6 - 2 * B)) * (1.f / 6.f);
            };
            return ev(x * invX) * ev(y * invY);
        }
    }
    return 0.0f;
}

}  // namespace orc
