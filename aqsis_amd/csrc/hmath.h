// Host-side single-precision arithmetic of the hider's front end.
//
// Every routine reproduces the operand order, float/double mix and
// comparison strictness of the aqsis code it names, because the dice sizes,
// bucket assignments and raster coordinates it feeds must be bit-identical
// to the reference's.  Compile with -ffp-contract=off.
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdint>

namespace rh {

// include/aqsis/math/math.h:47-70
template <typename T> inline long floor_l(T x) { return static_cast<long>(x) - (x < 0 && x != static_cast<long>(x)); }
template <typename T> inline long ceil_l(T x) { return static_cast<long>(x) + (x > 0 && x != static_cast<long>(x)); }
template <typename T> inline long round_l(T x) { return floor_l(x - 0.5) + 1; }
template <typename T> inline T mn(T a, T b) { return (a < b) ? a : b; }
template <typename T> inline T mx(T a, T b) { return (a < b) ? b : a; }
template <typename T> inline T clampv(T x, T lo, T hi) { return x < lo ? lo : (x > hi ? hi : x); }
inline float deg2rad(float a) { return (float)((double)a / 180.0 * M_PI); }

struct f3 {
    float x, y, z;
};
inline f3 mk(float x, float y, float z) { f3 r = {x, y, z}; return r; }
inline f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline f3 scale(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float len2(f3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline f3 crs(f3 a, f3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline f3 normalised(f3 a) {
    float l = std::sqrt(len2(a));
    if (l != 0) { a.x /= l; a.y /= l; a.z /= l; }
    return a;
}

struct box3 {
    f3 lo, hi;
};
inline void grow(box3& b, f3 v) {
    b.hi.x = mx(b.hi.x, v.x); b.hi.y = mx(b.hi.y, v.y); b.hi.z = mx(b.hi.z, v.z);
    b.lo.x = mn(b.lo.x, v.x); b.lo.y = mn(b.lo.y, v.y); b.lo.z = mn(b.lo.z, v.z);
}
inline box3 empty_box() {
    box3 b;
    b.lo = mk(FLT_MAX, FLT_MAX, FLT_MAX);
    b.hi = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    return b;
}

// 4x4 matrix with aqsis's element layout e[row][col] (points are row
// vectors) and its identity short-cuts (include/aqsis/math/matrix.h).
struct m44 {
    float e[4][4];
    bool id;
};
inline m44 m_identity() {
    m44 m;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) m.e[i][j] = i == j ? 1.0f : 0.0f;
    m.id = true;
    return m;
}
inline m44 m_diag(float a, float b, float c) {  // matrix.h:404-417
    m44 m = m_identity();
    if (a != 1.0f || b != 1.0f || c != 1.0f) {
        m.e[0][0] = a; m.e[1][1] = b; m.e[2][2] = c; m.e[3][3] = 1.0f; m.id = false;
    }
    return m;
}
inline m44 m_trans(float x, float y, float z) {  // matrix.h:420-432
    m44 m = m_identity();
    if (x != 0.0f || y != 0.0f || z != 0.0f) { m.id = false; m.e[3][0] = x; m.e[3][1] = y; m.e[3][2] = z; }
    return m;
}
// lhs * rhs as aqsis's operator* (matrix.cpp:150-228): r[i][j] = sum_k rhs[i][k]*lhs[k][j].
inline m44 m_mul(const m44& lhs, const m44& rhs) {
    if (lhs.id) return rhs;
    if (rhs.id) return lhs;
    m44 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.e[i][j] = rhs.e[i][0] * lhs.e[0][j] + rhs.e[i][1] * lhs.e[1][j] + rhs.e[i][2] * lhs.e[2][j] +
                        rhs.e[i][3] * lhs.e[3][j];
    r.id = false;
    return r;
}
// CqMatrix * CqVector3D (matrix.h:717-750)
inline f3 m_point(const m44& m, f3 v) {
    if (m.id) return v;
    float h = m.e[0][3] * v.x + m.e[1][3] * v.y + m.e[2][3] * v.z + m.e[3][3];
    f3 r;
    r.x = m.e[0][0] * v.x + m.e[1][0] * v.y + m.e[2][0] * v.z + m.e[3][0];
    r.y = m.e[0][1] * v.x + m.e[1][1] * v.y + m.e[2][1] * v.z + m.e[3][1];
    r.z = m.e[0][2] * v.x + m.e[1][2] * v.y + m.e[2][2] * v.z + m.e[3][2];
    if (h != 1) {
        float inv = 1 / h;
        r.x = r.x * inv; r.y = r.y * inv; r.z = r.z * inv;
    }
    return r;
}
// CqMatrix::Rotate on identity (matrix.cpp:96-135)
inline m44 m_rot(float angle, f3 axis) {
    m44 R = m_identity();
    if (angle != 0.0f && std::sqrt(len2(axis)) != 0.0f) {
        R.id = false;
        f3 a = normalised(axis);
        float s = (float)std::sin((double)angle), c = (float)std::cos((double)angle), t = 1.0f - c;
        R.e[0][0] = t * a.x * a.x + c;
        R.e[1][1] = t * a.y * a.y + c;
        R.e[2][2] = t * a.z * a.z + c;
        float txy = t * a.x * a.y, sz = s * a.z;
        R.e[0][1] = txy + sz; R.e[1][0] = txy - sz;
        float txz = t * a.x * a.z, sy = s * a.y;
        R.e[0][2] = txz - sy; R.e[2][0] = txz + sy;
        float tyz = t * a.y * a.z, sx = s * a.x;
        R.e[1][2] = tyz + sx; R.e[2][1] = tyz - sx;
    }
    return R;
}
// CqBound::Transform (bound.cpp:75-109)
inline box3 box_xform(const box3& b, const m44& m) {
    f3 c[8] = {b.lo, mk(b.hi.x, b.lo.y, b.lo.z), mk(b.lo.x, b.hi.y, b.lo.z), mk(b.lo.x, b.lo.y, b.hi.z),
               mk(b.hi.x, b.hi.y, b.lo.z), mk(b.lo.x, b.hi.y, b.hi.z), mk(b.hi.x, b.lo.y, b.hi.z), b.hi};
    box3 r = empty_box();
    for (int i = 0; i < 8; ++i) {
        f3 p = m_point(m, c[i]);
        if (p.x < r.lo.x) r.lo.x = p.x;
        if (p.y < r.lo.y) r.lo.y = p.y;
        if (p.z < r.lo.z) r.lo.z = p.z;
        if (p.x > r.hi.x) r.hi.x = p.x;
        if (p.y > r.hi.y) r.hi.y = p.y;
        if (p.z > r.hi.z) r.hi.z = p.z;
    }
    return r;
}

}  // namespace rh
