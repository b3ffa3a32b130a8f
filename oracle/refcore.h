// ORACLE -- test infrastructure only.  This is a CPU restatement of aqsis's
// hider path used to check the HIP hider; it is never linked into, loaded
// by, or called from the product library.  Only tests/, the smoke check and
// the CPU-baseline leg of bench.py use it.
//
// refcore.h: arithmetic building blocks, each following the reference file
// and line cited beside it.  All single-precision code must be compiled with
// -ffp-contract=off on x86-64 SSE (no FMA, no x87), like the reference's
// default g++ build.
#ifndef ORACLE_REFCORE_H
#define ORACLE_REFCORE_H

#include <cfloat>
#include <cmath>
#include <limits>
#include <cstdint>
#include <vector>
#include <algorithm>

namespace orc {

// ---------------------------------------------------------------------------
// Rounding helpers: include/aqsis/math/math.h:47-70.
template <typename T> inline long lfloor(const T x) {
    return static_cast<long>(x) - (x < 0 && x != static_cast<long>(x));
}
template <typename T> inline long lceil(const T x) {
    return static_cast<long>(x) + (x > 0 && x != static_cast<long>(x));
}
// lround(x) = lfloor(x - 0.5) + 1; the 0.5 literal is a double, so a float
// argument is promoted before the subtraction (math.h:67-70).
template <typename T> inline long lround(const T x) { return lfloor(x - 0.5) + 1; }

template <typename T> inline T tmin(const T a, const T b) { return (a < b) ? a : b; }  // math.h:120-125
template <typename T> inline T tmax(const T a, const T b) { return (a < b) ? b : a; }  // math.h:131-136
template <typename T> inline T tclamp(const T x, const T lo, const T hi) {              // math.h:101-106
    return x < lo ? lo : (x > hi ? hi : x);
}

// degToRad: a/180.0*M_PI evaluated in double, returned as float (math.h:140-143).
inline float degToRad(const float a) { return (float)((double)a / 180.0 * M_PI); }

// log2 = 1.4426950408889633 * std::log(x) with x float -> logf (math.h:157-161).
inline float aq_log2(const float x) { return (float)(1.4426950408889633 * (double)std::log(x)); }

// ---------------------------------------------------------------------------
// Vectors: include/aqsis/math/vector2d.h, vector3d.h, color.h.
struct V2 { float x, y; };
struct V3 { float x, y, z; };

inline V2 v2(float x, float y) { V2 r; r.x = x; r.y = y; return r; }
inline V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
inline V2 operator+(V2 a, V2 b) { return v2(a.x + b.x, a.y + b.y); }
inline V2 operator-(V2 a, V2 b) { return v2(a.x - b.x, a.y - b.y); }
inline V2 operator*(V2 a, float f) { return v2(a.x * f, a.y * f); }
inline V2 operator*(float f, V2 a) { return v2(f * a.x, f * a.y); }
inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator*(float f, V3 a) { return v3(f * a.x, f * a.y, f * a.z); }
inline V3 operator*(V3 a, float f) { return v3(a.x * f, a.y * f, a.z * f); }
// dot: vector3d.h:603-606
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// cross: vector3d.h:411-417
inline V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float mag2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }     // vector3d.h:345-348
inline float mag(V3 a) { return std::sqrt(mag2(a)); }                    // vector3d.h:350-353
inline V3 unit(V3 a) {                                                   // vector3d.h:355-365
    float l = mag(a);
    if (l != 0) { a.x /= l; a.y /= l; a.z /= l; }
    return a;
}
inline V3 vmin(V3 a, V3 b) { return v3(tmin(a.x, b.x), tmin(a.y, b.y), tmin(a.z, b.z)); }
inline V3 vmax(V3 a, V3 b) { return v3(tmax(a.x, b.x), tmax(a.y, b.y), tmax(a.z, b.z)); }
// 2D helpers: vector2d.h:372-385
inline float cross2(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }
inline float maxNorm(V2 v) { return tmax(std::fabs(v.x), std::fabs(v.y)); }
// CqVector2D::Magnitude2 (vector2d.h:132-138)
inline float mag2_2d(float x, float y) {
    if (y == 0.0) return x * x;
    else if (x == 0.0) return y * y;
    return (x * x) + (y * y);
}
// isClose (math.h:174-182)
inline bool isClose(float x1, float x2, float tol = 10 * std::numeric_limits<float>::epsilon()) {
    float d = std::fabs(x1 - x2);
    return d <= tol * std::fabs(x1) || d <= tol * std::fabs(x2);
}

// ---------------------------------------------------------------------------
// CqMatrix: include/aqsis/math/matrix.h, libs/math/matrix.cpp.  Element
// m[row][col]; a point transforms as a row vector (translation in row 3).
struct Mat {
    float m[4][4];
    bool ident;
};
inline Mat mat_identity() {
    Mat r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = (i == j) ? 1.0f : 0.0f;
    r.ident = true;
    return r;
}
// CqMatrix(xs, ys, zs): matrix.h:404-417
inline Mat mat_scale(float xs, float ys, float zs) {
    Mat r = mat_identity();
    if (xs != 1.0f || ys != 1.0f || zs != 1.0f) {
        r.m[0][0] = xs; r.m[1][1] = ys; r.m[2][2] = zs; r.m[3][3] = 1.0f;
        r.ident = false;
    }
    return r;
}
// CqMatrix(trans): matrix.h:420-432
inline Mat mat_translate(float x, float y, float z) {
    Mat r = mat_identity();
    if (x != 0.0f || y != 0.0f || z != 0.0f) {
        r.ident = false;
        r.m[3][0] = x; r.m[3][1] = y; r.m[3][2] = z;
    }
    return r;
}
// A * B == temp(A); temp *= B, and *= computes from[i][k]*A[k][j] summed
// left to right (matrix.cpp:150-228, matrix.h:636-641).
inline Mat mat_mul(const Mat& A, const Mat& from) {
    if (A.ident) return from;
    if (from.ident) return A;
    Mat r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.m[i][j] = from.m[i][0] * A.m[0][j] + from.m[i][1] * A.m[1][j] +
                        from.m[i][2] * A.m[2][j] + from.m[i][3] * A.m[3][j];
    r.ident = false;
    return r;
}
// CqMatrix::operator*(CqVector3D): matrix.h:717-750 (homogeneous divide by
// a correctly rounded reciprocal when h != 1).
inline V3 mat_vec3(const Mat& M, V3 v) {
    if (M.ident) return v;
    const float(*e)[4] = M.m;
    float h = e[0][3] * v.x + e[1][3] * v.y + e[2][3] * v.z + e[3][3];
    V3 r;
    r.x = e[0][0] * v.x + e[1][0] * v.y + e[2][0] * v.z + e[3][0];
    r.y = e[0][1] * v.x + e[1][1] * v.y + e[2][1] * v.z + e[3][1];
    r.z = e[0][2] * v.x + e[1][2] * v.y + e[2][2] * v.z + e[3][2];
    if (h != 1) {
        float invh = 1 / h;
        r.x = r.x * invh; r.y = r.y * invh; r.z = r.z * invh;
    }
    return r;
}
// CqMatrix::Rotate (matrix.cpp:96-135) applied to an identity matrix.
Mat mat_rotate(float angle, V3 axis);

// ---------------------------------------------------------------------------
// Global MT19937 stream of libs/math/random.cpp:90-224 (one shared state;
// constructor seeds are ignored, random.cpp:173-174).
struct Rng {
    uint32_t mt[624];
    int mti = 625;
    void reseed(uint32_t s);          // init_genrand, random.cpp:102-116
    uint32_t next();                  // genrand_int32, random.cpp:119-157
    // RandomFloat: u32 * (1.0/4294967424.0) in double, returned as float (random.cpp:194-206)
    float rfloat() { return (float)((double)next() * (1.0 / 4294967424.0)); }
    // RandomFloat(range) = range * RandomFloat() in float (random.cpp:211-214)
    float rfloat(float range) { return range * rfloat(); }
    // RandomInt(range) = lfloor((double)RandomFloat((float)range)) (random.cpp:186-190)
    uint32_t rint(uint32_t range) {
        double n = rfloat((float)range);
        return (uint32_t)lfloor(n);
    }
};

// ---------------------------------------------------------------------------
// CqMultiJitteredSampler (libs/core/multijitter.h:47-101, multijitter.cpp:83-222)
// and CqGridSampler (libs/core/grid.cpp:37-79).
struct SamplerTables {
    int nx = 0, ny = 0, n = 0, cache = 0;
    std::vector<V2> pos2d;        // cache * n
    std::vector<float> pos1d;     // cache * n
    std::vector<int> shuffled;    // cache * n
};
// Builds the 250 cached patterns from the stream and then reseeds it to 19.
void build_multijitter(Rng& rng, int nx, int ny, SamplerTables& t);
void build_grid_sampler(int nx, int ny, SamplerTables& t);

// ---------------------------------------------------------------------------
// Pixel filters: libs/core/filters.cpp:72-342.
float filter_eval(int filter, float x, float y, float xw, float yw);

}  // namespace orc

#endif
