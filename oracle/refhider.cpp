// ORACLE -- test infrastructure only.  CPU restatement of the aqsis hider
// path; see refhider.h.  Each function names the reference code it follows.
// Simplifications, all result-neutral for full-frame renders:
//  * pixels live in one extended-image array instead of per-bucket arrays
//    handed over through cache segments (bucketprocessor.cpp:259-355,
//    1645-1694): the reference's hand-over gives every extended pixel to the
//    first bucket (row-major) whose data region holds it, and the sample
//    region below is computed with the same rule;
//  * hit records keep the 7 floats the path writes (Ci, Oi, depth); the
//    coverage and alpha slots are rebuilt after filtering
//    (bucketprocessor.cpp:690-707);
//  * CSG, matte, AOVs, trim curves, LOD ranges, DoF and imager shaders are
//    not on this path.
#include "refhider.h"
#include "refcore.h"

#include <chrono>
#include <climits>
#include <limits>
#include <cstring>
#include <memory>
#include <stdexcept>

namespace orc {

// ---------------------------------------------------------------------------
// CqBound (libs/core/bound.h, bound.cpp)
struct Bound {
    V3 mn, mx;
};
static inline void encapsulate(Bound& b, V3 v) {  // bound.cpp:134-143
    b.mx.x = tmax(b.mx.x, v.x); b.mx.y = tmax(b.mx.y, v.y); b.mx.z = tmax(b.mx.z, v.z);
    b.mn.x = tmin(b.mn.x, v.x); b.mn.y = tmin(b.mn.y, v.y); b.mn.z = tmin(b.mn.z, v.z);
}
static Bound bound_transform(const Bound& b, const Mat& M) {  // bound.cpp:75-109
    V3 c[8];
    c[0] = b.mn;
    c[1] = v3(b.mx.x, b.mn.y, b.mn.z);
    c[2] = v3(b.mn.x, b.mx.y, b.mn.z);
    c[3] = v3(b.mn.x, b.mn.y, b.mx.z);
    c[4] = v3(b.mx.x, b.mx.y, b.mn.z);
    c[5] = v3(b.mn.x, b.mx.y, b.mx.z);
    c[6] = v3(b.mx.x, b.mn.y, b.mx.z);
    c[7] = b.mx;
    Bound r;
    r.mn = v3(FLT_MAX, FLT_MAX, FLT_MAX);
    r.mx = v3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    for (int i = 0; i < 8; ++i) {
        V3 p = mat_vec3(M, c[i]);
        if (p.x < r.mn.x) r.mn.x = p.x;
        if (p.y < r.mn.y) r.mn.y = p.y;
        if (p.z < r.mn.z) r.mn.z = p.z;
        if (p.x > r.mx.x) r.mx.x = p.x;
        if (p.y > r.mx.y) r.mx.y = p.y;
        if (p.z > r.mx.z) r.mx.z = p.z;
    }
    return r;
}

// ---------------------------------------------------------------------------
// Depth of field: CqRenderer::SetDepthOfFieldData / SetDepthOfFieldScale
// (renderer.h:368-396, options.cpp:160-183).
struct Dof {
    bool on = false;
    float mult = 0, oneOverF = 0;  // m_DofMultiplier, m_OneOverFocalDistance
    V2 scale = {0, 0};             // m_DepthOfFieldScale
};
// GetCircleOfConfusion (renderer.h:401-406): the raster radius of the
// circle of confusion at a camera depth.
static V2 circle_of_confusion(const Dof& d, float depth) {
    float c = d.mult * std::fabs(1.0f / depth - d.oneOverF);
    return v2(d.scale.x * c, d.scale.y * c);
}
// MinCoCForBound (renderer.cpp:1602-1617)
static float min_coc_for_bound(const Dof& d, float z1, float z2) {
    float focalDist = 1 / d.oneOverF;
    if ((z1 - focalDist) * (z2 - focalDist) < 0) return 0;
    float minBlur = tmin(std::fabs(1 / z1 - d.oneOverF), std::fabs(1 / z2 - d.oneOverF));
    return d.mult * tmin(d.scale.x, d.scale.y) * minBlur;
}
// CqImagePixel::projectToCircle (imagepixel.h:429-436): the unit square to
// the unit disk, radially.
static V2 project_to_circle(V2 pos) {
    float r;
    if (pos.y == 0.0f) r = std::fabs(pos.x);
    else if (pos.x == 0.0f) r = std::fabs(pos.y);
    else r = std::sqrt((pos.x * pos.x) + (pos.y * pos.y));  // CqVector2D::Magnitude
    if (r == 0.0f) return v2(0, 0);
    float adj = tmax(std::fabs(pos.x), std::fabs(pos.y)) / r;
    return v2(adj * pos.x, adj * pos.y);
}

// ---------------------------------------------------------------------------
// Frame options: SqOptionCache (optioncache.cpp:48-117), camera
// (options.cpp:40-160, ri.cpp:586-617), crop window (renderer.cpp:1619-1626).
struct Frame {
    rh_options o;
    int cropXMin, cropXMax, cropYMin, cropYMax;
    int shiftX, shiftY;          // DiscreteShift = lfloor(filtsize/2)
    float xfwo2, yfwo2;          // ceil(filtsize)*0.5
    int nx, ny, nsamp;
    int cXBuckets, cYBuckets;
    int bxMin, byMin, bxMax, byMax;  // bucket region
    Mat camToRaster;
    Mat diceCoords;              // camToRaster with the z column zeroed
    int extX0, extY0, extW, extH;  // extended (sampled) image
    std::vector<float> filterValues;  // InitialiseFilterValues
    Dof dof;                     // lens (CqRenderer's depth-of-field state)
    std::vector<Bound> dofBounds;  // CalculateDofBounds: one per lens cell
    // isCullable without CSG (bucketprocessor.cpp:1076-1081): false with a z
    // display and the max or average depth filter
    bool cullableHits = true;
    // StoreSample's midpoint case (bucketprocessor.cpp:1502-1525)
    bool midpointStore = false;
};

static void setup_frame(const rh_options& o, Frame& f) {
    f.o = o;
    const int xr = o.xres, yr = o.yres;
    // initialiseCropWindow: lceil(res * crop) clamped (renderer.cpp:1619-1626)
    f.cropXMin = (int)tclamp<long>(lceil((float)xr * o.crop[0]), 0, xr);
    f.cropXMax = (int)tclamp<long>(lceil((float)xr * o.crop[1]), 0, xr);
    f.cropYMin = (int)tclamp<long>(lceil((float)yr * o.crop[2]), 0, yr);
    f.cropYMax = (int)tclamp<long>(lceil((float)yr * o.crop[3]), 0, yr);
    f.shiftX = (int)lfloor(o.filter_xwidth / 2.0f);
    f.shiftY = (int)lfloor(o.filter_ywidth / 2.0f);
    f.xfwo2 = std::ceil(o.filter_xwidth) * 0.5f;
    f.yfwo2 = std::ceil(o.filter_ywidth) * 0.5f;
    f.nx = o.xsamples; f.ny = o.ysamples; f.nsamp = f.nx * f.ny;
    f.cullableHits = !(o.display_z && (o.depth_filter == RH_DEPTH_MAX || o.depth_filter == RH_DEPTH_AVERAGE));
    f.midpointStore = o.display_z && o.depth_filter == RH_DEPTH_MIDPOINT;
    // SetImage (imagebuffer.cpp:174-228)
    f.cXBuckets = (xr - 1) / o.bucket_x + 1;
    f.cYBuckets = (yr - 1) / o.bucket_y + 1;
    f.bxMin = f.cropXMin / o.bucket_x;
    f.byMin = f.cropYMin / o.bucket_y;
    f.bxMax = (f.cropXMax - 1) / o.bucket_x + 1;
    f.byMax = (f.cropYMax - 1) / o.bucket_y + 1;

    // Frame aspect ratio and default screen window (ri.cpp:590-617).
    float PAR = o.pixel_aspect;
    float resH = (float)xr, resV = (float)yr;
    float FAR = (resH * PAR) / resV;
    float l, r, t, b;
    if (FAR >= 1.0) { l = -FAR; r = +FAR; t = +1; b = -1; }
    else { l = -1; r = +1; t = (float)(+1.0 / FAR); b = (float)(-1.0 / FAR); }
    // InitialiseCamera, perspective (options.cpp:40-160).
    float n = o.clip_near, fcl = o.clip_far;
    float s = (float)std::tan((double)(degToRad(o.fov) / 2.0f));
    float a = fcl / (fcl - n);
    Mat camToImage = mat_identity();
    camToImage.ident = false;
    camToImage.m[0][0] = 1.0f / s;
    camToImage.m[1][1] = 1.0f / s;
    camToImage.m[2][2] = a;
    camToImage.m[3][2] = -n * a;
    camToImage.m[2][3] = 1.0f;
    camToImage.m[3][3] = 0.0f;
    Mat imageToScreen = mat_identity();
    imageToScreen.ident = false;
    imageToScreen.m[0][0] = 2.0f / (r - l);
    imageToScreen.m[1][1] = 2.0f / (t - b);
    imageToScreen.m[3][0] = (-(r + l)) / (r - l);
    imageToScreen.m[3][1] = (-(t + b)) / (t - b);
    Mat preProj = mat_identity();
    Mat camToScreen = mat_mul(mat_mul(imageToScreen, preProj), camToImage);
    float FrameX = (FAR >= 1.0) ? (float)xr : ((float)yr * FAR) / PAR;
    float FrameY = (FAR < 1.0) ? (float)yr : ((float)xr * PAR) / FAR;
    Mat T = mat_translate(1.0f, 1.0f, 0.0f);
    Mat S = mat_scale(0.5f, 0.5f, 1.0f);
    Mat S2 = mat_scale(FrameX, FrameY, 1.0f);
    Mat S3 = mat_scale(1.0f, -1.0f, 1.0f);
    Mat screenToNDC = mat_mul(mat_mul(S, T), S3);
    Mat ndcToRaster = S2;
    Mat worldToCamera = mat_identity();
    Mat raster = mat_mul(mat_mul(mat_mul(ndcToRaster, screenToNDC), camToScreen), worldToCamera);
    // matSpaceToSpace("camera","raster") = matWorldToRaster * identity (renderer.cpp:838-889)
    f.camToRaster = mat_mul(raster, mat_identity());
    // RenderSurface: zero the z output column (bucketprocessor.cpp:1008-1012)
    f.diceCoords = f.camToRaster;
    f.diceCoords.m[0][2] = f.diceCoords.m[1][2] = 0;
    f.diceCoords.m[2][2] = f.diceCoords.m[3][2] = 0;

    // The CoC scale: matVSpaceToSpace("camera", "raster") -- the same product
    // with its translation and projective column cleared (renderer.cpp:
    // 897-957) -- applied to (1,1,-1) and (0,0,-1) (options.cpp:160-171).
    {
        Mat dofm = f.camToRaster;
        dofm.m[3][0] = dofm.m[3][1] = dofm.m[3][2] = dofm.m[0][3] = dofm.m[1][3] = dofm.m[2][3] = 0.0f;
        dofm.m[3][3] = 1.0f;
        V3 dofe = mat_vec3(dofm, v3(1, 1, -1));
        V3 dofc = mat_vec3(dofm, v3(0, 0, -1));
        f.dof.scale = v2(std::fabs(dofe.x - dofc.x), std::fabs(dofe.y - dofc.y));
        // SetDepthOfFieldData (renderer.h:368-377), double intermediates
        const float fstop = o.dof[0], focalLength = o.dof[1], focalDistance = o.dof[2];
        f.dof.on = fstop < FLT_MAX;
        if (f.dof.on) {
            float lensDiameter = focalLength / fstop;
            f.dof.mult = (float)(0.5 * lensDiameter * focalDistance / (focalDistance + lensDiameter));
            f.dof.oneOverF = (float)(1.0 / focalDistance);
        }
    }
    // CalculateDofBounds (bucketprocessor.cpp:858-911): the bound of the lens
    // offsets of each of the xs * ys cells of [-1,1]^2 after projectToCircle.
    if (f.dof.on) {
        f.dofBounds.assign(f.nsamp, Bound());
        float dx = (float)(2.0 / f.nx), dy = (float)(2.0 / f.ny);
        float minX = -1.0f, minY = -1.0f;
        int which = 0;
        for (int j = 0; j < f.ny; ++j) {
            for (int i = 0; i < f.nx; ++i) {
                V2 tl = project_to_circle(v2(minX, minY)), tr = project_to_circle(v2(minX + dx, minY));
                V2 bl = project_to_circle(v2(minX, minY + dy)), br = project_to_circle(v2(minX + dx, minY + dy));
                if ((tl.y > 0.0f && bl.y < 0.0f) || (tl.y < 0.0f && bl.y > 0.0f)) {
                    tl.x = minX; bl.x = minX; tr.x = minX + dx; br.x = minX + dx;
                }
                if ((tl.x > 0.0f && tr.x < 0.0f) || (tl.x < 0.0f && tr.x > 0.0f)) {
                    tl.y = minY; bl.y = minY + dy; tr.y = minY; br.y = minY + dy;
                }
                Bound& b = f.dofBounds[which];
                b.mn = v3(tl.x, tl.y, 0.0f);
                b.mx = v3(tl.x, tl.y, 0.0f);
                encapsulate(b, v3(tr.x, tr.y, 0.0f));
                encapsulate(b, v3(bl.x, bl.y, 0.0f));
                encapsulate(b, v3(br.x, br.y, 0.0f));
                ++which;
                minX += dx;
            }
            minX = -1.0f;
            minY += dy;
        }
    }

    f.extX0 = f.cropXMin - f.shiftX;
    f.extY0 = f.cropYMin - f.shiftY;
    f.extW = (f.cropXMax - f.cropXMin) + 2 * f.shiftX;
    f.extH = (f.cropYMax - f.cropYMin) + 2 * f.shiftY;

    // InitialiseFilterValues (bucketprocessor.cpp:811-856)
    int xmax = f.shiftX, ymax = f.shiftY;
    int numPixels = (2 * xmax + 1) * (2 * ymax + 1);
    f.filterValues.assign((size_t)numPixels * f.nsamp, 0.0f);
    float cw = std::ceil(o.filter_xwidth), ch = std::ceil(o.filter_ywidth);
    for (int py = -ymax; py <= ymax; ++py)
        for (int px = -xmax; px <= xmax; ++px) {
            int idx = ((py + ymax) * (2 * xmax + 1) + px + xmax) * f.nsamp;
            for (int sy = 0; sy < f.ny; ++sy)
                for (int sx = 0; sx < f.nx; ++sx, ++idx) {
                    float fx = (sx + 0.5f) / f.nx + px - 0.5f;
                    float fy = (sy + 0.5f) / f.ny + py - 0.5f;
                    float w = 0;
                    if (fx >= -f.xfwo2 && fy >= -f.yfwo2 && fx <= f.xfwo2 && fy <= f.yfwo2)
                        w = filter_eval(o.filter, fx, fy, cw, ch);
                    f.filterValues[idx] = w;
                }
        }
}

// ---------------------------------------------------------------------------
// Surfaces: CqSurface (geometry/surface.{h,cpp}), CqSphere / CqQuadric
// (geometry/quadrics.cpp), CqSurfacePatchBicubic / Bilinear (patch.cpp).
enum { SplitDir_U = 0, SplitDir_V = 1 };

struct Surf {
    int type = 0;
    int prim = 0;
    int splitCount = 0;
    int splitDir = SplitDir_U;
    bool fDiceable = true;
    bool fDiscard = false;
    int uDice = 0, vDice = 0;
    bool cachedBound = false;
    Bound rasterBound;
    // sphere
    float radius = 1, phiMin = 0, phiMax = 0, thetaMin = 0, thetaMax = 360;
    Mat matTx;
    // patches: camera-space control points
    V3 cp[16];
    // Deformation blur (CqDeformingSurface, geometry/surface.h:617-720): the
    // fields above are the deforming surface's own state plus its key-0
    // surface; key 1's geometry follows.  The key surfaces keep their own
    // m_fDiceable / m_fDiscard, which nothing clears.
    bool motion = false;
    float keyTime[2] = {0.0f, 0.0f};
    V3 cp1[16];
    Mat matTx1;
};

// The key-1 surface of a deforming surface.
static Surf key1_of(const Surf& s) {
    Surf k = s;
    k.motion = false;
    for (int i = 0; i < 16; ++i) k.cp[i] = s.cp1[i];
    k.matTx = s.matTx1;
    k.fDiceable = true;
    k.fDiscard = false;
    return k;
}

struct Grid {
    int cu = 0, cv = 0;
    std::vector<V3> P;  // camera space after dicing; raster x,y + camera z after bust
    float Ci[3], Oi[3];
    bool constantShading = false;   // ShadingInterpolation "constant"
    uint32_t matteFlag = 0;         // CacheGridInfo (micropolygon.cpp:49-55)
    uint32_t mpBase = 0;
    int refs = 0;
};

struct MP {
    Grid* g = nullptr;
    int index = 0;
    uint32_t id = 0;
    int indexCode = 0;
    Bound bound;
    int refs = 0;
    // CqMicroPolygonMotion (micropolygon.h:872-965): two keys of the four
    // raster vertices (grid order 0, 1, cu+1, cu+2), their times and bounds,
    // and the sub-bound list built on first use.
    bool motion = false;
    V3 K[2][4];
    float T[2] = {0.0f, 0.0f};
    Bound keyBound[2];
    std::vector<Bound> subBound;
    std::vector<float> subTime;
};
static const int Degeneracy_Mask = 0x8000000;  // micropolygon.h:791

// Circle (quadrics.cpp:1360-1400).  Unqualified cos/sin of float arguments
// resolve to the C library's double functions.
static void circle(V3 O, V3 X, V3 Y, float r, float as, float ae, std::vector<V3>& pts);
static bool intersect_line(V3 P1, V3 T1, V3 P2, V3 T2, V3& P);
static void project_to_line(V3 S, V3 Trj, V3 pnt, V3& p);

static void circle(V3 O, V3 X, V3 Y, float r, float as, float ae, std::vector<V3>& points) {
    const double RI_PI = 3.14159265358979323846;
    while (ae < as) ae = (float)(ae + 2 * RI_PI);
    float theta = ae - as;
    unsigned narcs = 4;
    float dtheta = theta / static_cast<float>(narcs);
    unsigned n = 2 * narcs + 1;
    V3 P0, T0, P2, T2, P1 = v3(0, 0, 0);
    P0 = (O + (float)(r * std::cos((double)as)) * X) + (float)(r * std::sin((double)as)) * Y;
    T0 = (float)(-std::sin((double)as)) * X + (float)std::cos((double)as) * Y;
    points.assign(n, v3(0, 0, 0));
    points[0] = P0;
    unsigned index = 0;
    float angle = as;
    for (unsigned i = 1; i <= narcs; ++i) {
        angle += dtheta;
        P2 = (O + (float)(r * std::cos((double)angle)) * X) + (float)(r * std::sin((double)angle)) * Y;
        points[index + 2] = P2;
        T2 = (float)(-std::sin((double)angle)) * X + (float)std::cos((double)angle) * Y;
        intersect_line(P0, T0, P2, T2, P1);
        points[index + 1] = P1;
        index += 2;
        if (i < narcs) { P0 = P2; T0 = T2; }
    }
}

// IntersectLine (quadrics.cpp:1527-1541), including its P1 - T2 term.
static bool intersect_line(V3 P1, V3 T1, V3 P2, V3 T2, V3& P) {
    V3 px = cross(T1, P1 - T2);
    V3 v = cross(px, T1);
    float t = dot(P1 - P2, v);
    float vw = dot(v, T2);
    if ((double)(vw * vw) < 1.0e-07) return false;
    t /= vw;
    (void)t;
    P = P2 + (dot(P1 - P2, v) / vw) * T2;
    return true;
}

// ProjectToLine (quadrics.cpp:1548-1556)
static void project_to_line(V3 S, V3 Trj, V3 pnt, V3& p) {
    V3 a = pnt - S;
    float denom = mag2(Trj);
    float fraction = (denom == 0.0) ? 0.0f : dot(Trj, a) / denom;
    p = fraction * Trj;
    p = p + S;
}

// RevolveForBound (quadrics.cpp:1404-1474)
static Bound revolve_for_bound(const std::vector<V3>& profile, V3 S, V3 Tvec, float theta) {
    const double RI_PI = 3.14159265358979323846;
    Bound bound;
    bound.mn = v3(FLT_MAX, FLT_MAX, FLT_MAX);
    bound.mx = v3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    if (std::fabs(theta) > 2.0 * RI_PI) theta = (theta < 0) ? (float)(-(2.0 * RI_PI)) : (float)(2.0 * RI_PI);
    unsigned narcs = 4;
    float dtheta = theta / static_cast<float>(narcs);
    std::vector<float> cosines(narcs + 1), sines(narcs + 1);
    for (unsigned i = 1; i <= narcs; ++i) {
        float angle = dtheta * static_cast<float>(i);
        cosines[i] = (float)std::cos((double)angle);
        sines[i] = (float)std::sin((double)angle);
    }
    V3 P0, T0, P2, T2, P1 = v3(0, 0, 0);
    for (size_t j = 0; j < profile.size(); ++j) {
        V3 O;
        V3 pj = profile[j];
        project_to_line(S, Tvec, pj, O);
        V3 X = pj - O;
        float r = mag(X);
        if ((double)r < 1e-7) { encapsulate(bound, O); continue; }
        X = unit(X);
        V3 Y = cross(Tvec, X);
        Y = unit(Y);
        P0 = profile[j];
        encapsulate(bound, P0);
        T0 = Y;
        for (unsigned i = 1; i <= narcs; ++i) {
            P2 = (O + (r * cosines[i]) * X) + (r * sines[i]) * Y;
            encapsulate(bound, P2);
            T2 = (-sines[i]) * X + cosines[i] * Y;
            intersect_line(P0, T0, P2, T2, P1);
            encapsulate(bound, P1);
            if (i < narcs) { P0 = P2; T0 = T2; }
        }
    }
    return bound;
}

// CqSphere::Bound (quadrics.cpp:391-406)
static Bound sphere_bound(const Surf& s) {
    std::vector<V3> curve;
    V3 vA = v3(0, 0, 0), vC = v3(0, 0, 1);
    V3 vB = v3(1, 0, 0);
    circle(vA, vB, vC, s.radius, tmin(s.phiMin, s.phiMax), tmax(s.phiMin, s.phiMax), curve);
    Mat matRot = mat_rotate(degToRad(s.thetaMin), vC);
    for (auto& p : curve) p = mat_vec3(matRot, p);
    Bound B = revolve_for_bound(curve, vA, vC, degToRad(s.thetaMax - s.thetaMin));
    return bound_transform(B, s.matTx);
}

// Patch bounds: control-point boxes (patch.cpp:172-197, 622-648)
static Bound patch_bound(const Surf& s, int n) {
    V3 a = v3(FLT_MAX, FLT_MAX, FLT_MAX), b = v3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    for (int i = 0; i < n; ++i) {
        V3 v = s.cp[i];
        if (v.x < a.x) a.x = v.x;
        if (v.y < a.y) a.y = v.y;
        if (v.x > b.x) b.x = v.x;
        if (v.y > b.y) b.y = v.y;
        if (v.z < a.z) a.z = v.z;
        if (v.z > b.z) b.z = v.z;
    }
    Bound r; r.mn = a; r.mx = b;
    return r;
}

static Bound surf_bound_static(const Surf& s) {
    if (s.type == RH_PRIM_SPHERE) return sphere_bound(s);
    return patch_bound(s, s.type == RH_PRIM_BICUBIC ? 16 : 4);
}

// CqDeformingSurface::Bound (surface.h:629-640): the keys' bounds encapsulated.
static Bound surf_bound(const Surf& s) {
    if (!s.motion) return surf_bound_static(s);
    Bound b;
    b.mn = v3(FLT_MAX, FLT_MAX, FLT_MAX);
    b.mx = v3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    Bound k0 = surf_bound_static(s), k1 = surf_bound_static(key1_of(s));
    for (const Bound* B : {&k0, &k1}) {
        b.mx = v3(tmax(b.mx.x, B->mx.x), tmax(b.mx.y, B->mx.y), tmax(b.mx.z, B->mx.z));
        b.mn = v3(tmin(b.mn.x, B->mn.x), tmin(b.mn.y, B->mn.y), tmin(b.mn.z, B->mn.z));
    }
    return b;
}

// sinCosGrid (quadrics.cpp:1493-1518): float endpoints, float step, double
// recurrence, float results.
static void sin_cos_grid(float t0, float t1, int numSteps, float* sint, float* cost) {
    double prevCos = std::cos((double)t0);
    double prevSin = std::sin((double)t0);
    double dt = (double)((t1 - t0) / (numSteps - 1));
    double cosDt = std::cos(dt);
    double sinDt = std::sin(dt);
    cost[0] = (float)prevCos;
    sint[0] = (float)prevSin;
    for (int i = 1; i < numSteps; ++i) {
        double currCos = cosDt * prevCos - sinDt * prevSin;
        double currSin = sinDt * prevCos + cosDt * prevSin;
        cost[i] = (float)currCos;
        sint[i] = (float)currSin;
        prevCos = currCos;
        prevSin = currSin;
    }
}

// CqSphere::DicePoints (quadrics.cpp:457-492) in object space.
static void sphere_dice_points(const Surf& s, int uDice, int vDice, std::vector<V3>& pts) {
    int thetaRes = uDice + 1, phiRes = vDice + 1;
    std::vector<float> sinT(thetaRes), cosT(thetaRes), sinP(phiRes), cosP(phiRes);
    sin_cos_grid(degToRad(s.thetaMin), degToRad(s.thetaMax), thetaRes, sinT.data(), cosT.data());
    sin_cos_grid(s.phiMin, s.phiMax, phiRes, sinP.data(), cosP.data());
    pts.resize((size_t)thetaRes * phiRes);
    for (int v = 0; v <= vDice; ++v)
        for (int u = 0; u <= uDice; ++u) {
            V3 unitP = v3(cosT[u] * cosP[v], sinT[u] * cosP[v], sinP[v]);
            pts[(size_t)v * (uDice + 1) + u] = s.radius * unitP;
        }
}

// CqSurface::AdjustedShadingRate (geometry/surface.cpp:688-724) without
// the motion factor: with depth of field the rate grows with the least
// circle of confusion over the surface's cached raster bound (its camera z
// range), GeometricFocusFactor times 0.025 of its area.
static thread_local const Dof* g_dof = nullptr;  // the frame this thread renders
static float adjusted_shading_rate(const Surf& s, const rh_prim& p) {
    float shadingRate = p.shading_rate;
    if (g_dof && g_dof->on) {
        const float focusFactor = p.focus_factor;
        const float minCoC = min_coc_for_bound(*g_dof, s.rasterBound.mn.z, s.rasterBound.mx.z);
        const float areaRatio = 0.025f;
        shadingRate *= tmax(1.0f, areaRatio * focusFactor * minCoC * minCoC);
    }
    return shadingRate;
}

// CqQuadric::EstimateGridSize + Diceable (quadrics.cpp:259-346)
static bool sphere_diceable(Surf& s, const Mat& matCtoR, const rh_prim& prim) {
    if (!s.fDiceable) return false;
    const int EST = 8;  // ESTIMATEGRIDSIZE, quadrics.h:37
    float maxusize = 0, maxvsize = 0;
    Mat matTx = mat_mul(matCtoR, s.matTx);
    std::vector<V3> grid;
    sphere_dice_points(s, EST, EST, grid);
    V3 p, pum1 = v3(0, 0, 0), pvm1[EST];
    for (int i = 0; i < EST; ++i) pvm1[i] = v3(0, 0, 0);
    for (int v = 0; v <= EST; ++v)
        for (int u = 0; u <= EST; ++u) {
            p = mat_vec3(matTx, grid[(size_t)v * (EST + 1) + u]);
            if (v >= 1 && u >= 1) {
                float udist = mag2(p - pum1);
                float vdist = mag2(pvm1[u - 1] - pum1);
                maxusize = tmax(maxusize, udist);
                maxvsize = tmax(maxvsize, vdist);
            }
            if (u >= 1) pvm1[u - 1] = pum1;
            pum1 = p;
        }
    maxusize = std::sqrt(maxusize);
    maxvsize = std::sqrt(maxvsize);
    float sqrtShadingRate = std::sqrt(adjusted_shading_rate(s, prim));
    s.uDice = (int)lceil(EST * maxusize / sqrtShadingRate);
    s.vDice = (int)lceil(EST * maxvsize / sqrtShadingRate);
    unsigned long toomuch = (unsigned long)s.uDice * s.vDice;
    s.splitDir = (s.uDice > s.vDice) ? SplitDir_U : SplitDir_V;
    float gs = 16.0f;
    if (toomuch > 10000) return false;
    if (s.uDice > gs) return false;
    if (s.vDice > gs) return false;
    return true;
}

// CqSurfacePatchBicubic::Diceable (patch.cpp:300-379)
static bool bicubic_diceable(Surf& s, const Mat& matCtoR, const rh_prim& prim) {
    if (!s.fDiceable) return false;
    V3 hull[16];
    for (int i = 0; i < 16; ++i) hull[i] = mat_vec3(matCtoR, s.cp[i]);
    float uLen = 0, vLen = 0;
    for (int u = 0; u < 16; u += 4) {
        V3 V1 = hull[u + 1] - hull[u], V2 = hull[u + 2] - hull[u + 1], V3_ = hull[u + 3] - hull[u + 2];
        if (mag2(V1) > uLen) uLen = mag2(V1);
        if (mag2(V2) > uLen) uLen = mag2(V2);
        if (mag2(V3_) > uLen) uLen = mag2(V3_);
    }
    for (int v = 0; v < 4; ++v) {
        V3 V1 = hull[v + 4] - hull[v], V2 = hull[v + 8] - hull[v + 4], V3_ = hull[v + 12] - hull[v + 8];
        if (mag2(V1) > vLen) vLen = mag2(V1);
        if (mag2(V2) > vLen) vLen = mag2(V2);
        if (mag2(V3_) > vLen) vLen = mag2(V3_);
    }
    float shadingRate = adjusted_shading_rate(s, prim);
    uLen = std::sqrt(uLen / shadingRate);
    vLen = std::sqrt(vLen / shadingRate);
    s.splitDir = (uLen > vLen) ? SplitDir_U : SplitDir_V;
    uLen *= 3;
    vLen *= 3;
    s.uDice = (int)tmax<long>(lround(uLen), 1);
    s.vDice = (int)tmax<long>(lround(vLen), 1);
    if (uLen < FLT_EPSILON || vLen < FLT_EPSILON) { s.fDiscard = true; return false; }
    float gs = 16.0f, gs2 = gs * gs;
    if (s.uDice > gs2 || s.vDice > gs2 || (s.uDice * s.vDice) > gs2) return false;
    return true;
}

// CqSurfacePatchBilinear::Diceable (patch.cpp:671-736)
static bool bilinear_diceable(Surf& s, const Mat& matCtoR, const rh_prim& prim) {
    if (!s.fDiceable) return false;
    V3 hull[4];
    for (int i = 0; i < 4; ++i) hull[i] = mat_vec3(matCtoR, s.cp[i]);
    V3 V1 = hull[1] - hull[0], V2 = hull[3] - hull[2];
    float uLen = (mag2(V1) > mag2(V2)) ? mag2(V1) : mag2(V2);
    V1 = hull[2] - hull[0]; V2 = hull[3] - hull[1];
    float vLen = (mag2(V1) > mag2(V2)) ? mag2(V1) : mag2(V2);
    float shadingRate = adjusted_shading_rate(s, prim);
    uLen = std::sqrt(uLen / shadingRate);
    vLen = std::sqrt(vLen / shadingRate);
    s.splitDir = (uLen > vLen) ? SplitDir_U : SplitDir_V;
    uLen = (float)tmax<long>(lround(uLen), 1);
    vLen = (float)tmax<long>(lround(vLen), 1);
    s.uDice = (int)uLen;
    s.vDice = (int)vLen;
    if (uLen < FLT_EPSILON || vLen < FLT_EPSILON) { s.fDiscard = true; return false; }
    float gs = 16.0f, gs2 = gs * gs;
    if (s.uDice > gs2 || s.vDice > gs2 || (s.uDice * s.vDice) > gs2) return false;
    return true;
}

static bool surf_diceable_key(Surf& s, const Mat& dc, const rh_prim& prim) {
    switch (s.type) {
        case RH_PRIM_SPHERE: return sphere_diceable(s, dc, prim);
        case RH_PRIM_BICUBIC: return bicubic_diceable(s, dc, prim);
        default: return bilinear_diceable(s, dc, prim);
    }
}

// CqDeformingSurface::Diceable (surface.h:698-707): key 0 decides (with its
// own m_fDiceable) and its dice sizes and split direction are copied to the
// other keys (CopySplitInfo, surface.h:163-168); a discard decided by key 0
// stays with key 0.
static bool surf_diceable(Surf& s, const Mat& dc, const rh_prim& prim) {
    if (!s.motion) return surf_diceable_key(s, dc, prim);
    const bool outerDiceable = s.fDiceable, outerDiscard = s.fDiscard;
    s.fDiceable = true;
    bool f = surf_diceable_key(s, dc, prim);
    s.fDiceable = outerDiceable;
    s.fDiscard = outerDiscard;
    return f;
}

// bicubicPatchNatSubdiv (patch.cpp:87-129) on true (dehomogenised) values;
// the reference's homogeneous weights are powers of two, so scaling by them
// is exact and the two forms round identically.
static void bicubic_subdiv(const V3* P, V3* R1, V3* R2, bool u) {
    if (u) {
        for (int iv = 0; iv < 4; ++iv) {
            int o = iv * 4;
            R1[o + 0] = P[o + 0];
            R1[o + 1] = (P[o + 0] + P[o + 1]) * 0.5f;
            R1[o + 2] = R1[o + 1] * 0.5f + (P[o + 1] + P[o + 2]) * 0.25f;
            R2[o + 3] = P[o + 3];
            R2[o + 2] = (P[o + 2] + P[o + 3]) * 0.5f;
            R2[o + 1] = R2[o + 2] * 0.5f + (P[o + 1] + P[o + 2]) * 0.25f;
            R1[o + 3] = (R1[o + 2] + R2[o + 1]) * 0.5f;
            R2[o + 0] = R1[o + 3];
        }
    } else {
        for (int iu = 0; iu < 4; ++iu) {
            R1[0 + iu] = P[0 + iu];
            R1[4 + iu] = (P[0 + iu] + P[4 + iu]) * 0.5f;
            R1[8 + iu] = R1[4 + iu] * 0.5f + (P[4 + iu] + P[8 + iu]) * 0.25f;
            R2[12 + iu] = P[12 + iu];
            R2[8 + iu] = (P[8 + iu] + P[12 + iu]) * 0.5f;
            R2[4 + iu] = R2[8 + iu] * 0.5f + (P[4 + iu] + P[8 + iu]) * 0.25f;
            R1[12 + iu] = (R1[8 + iu] + R2[4 + iu]) * 0.5f;
            R2[0 + iu] = R1[12 + iu];
        }
    }
}

// surfaceNaturalSubdivide (surface.cpp:301-327) for the bilinear corners.
static void bilinear_subdiv(const V3* P, V3* R1, V3* R2, bool u) {
    for (int i = 0; i < 4; ++i) { R1[i] = P[i]; R2[i] = P[i]; }
    if (u) {
        R2[1] = P[1]; R2[3] = P[3];
        R1[1] = R2[0] = (P[0] + P[1]) * 0.5f;
        R1[3] = R2[2] = (P[2] + P[3]) * 0.5f;
    } else {
        R2[2] = P[2]; R2[3] = P[3];
        R1[2] = R2[0] = (P[0] + P[2]) * 0.5f;
        R1[3] = R2[1] = (P[1] + P[3]) * 0.5f;
    }
}

// CqSurface::Split (surface.cpp:529-577) / CqSurfacePatchBilinear::Split
// (patch.cpp:748-...) without phantom vertices.
static void surf_split_motion(const Surf& s, std::vector<Surf*>& out);

static void surf_split(const Surf& s, std::vector<Surf*>& out) {
    if (s.motion) {
        surf_split_motion(s, out);
        return;
    }
    Surf* a = new Surf(s);
    Surf* b = new Surf(s);
    bool u = s.splitDir == SplitDir_U;
    if (s.type == RH_PRIM_SPHERE) {
        // CqSphere::PreSubdivide (quadrics.cpp:413-451)
        float phicent = (float)((s.phiMin + s.phiMax) * 0.5);
        float arccent = (float)((s.thetaMin + s.thetaMax) * 0.5);
        if (u) {
            a->thetaMax = arccent; b->thetaMin = arccent;
        } else {
            a->phiMax = phicent; b->phiMin = phicent;
        }
    } else if (s.type == RH_PRIM_BICUBIC) {
        bicubic_subdiv(s.cp, a->cp, b->cp, u);
    } else {
        bilinear_subdiv(s.cp, a->cp, b->cp, u);
    }
    for (Surf* c : {a, b}) {
        c->splitDir = u ? SplitDir_V : SplitDir_U;
        c->splitCount = s.splitCount + 1;
        c->fDiceable = true;
        c->fDiscard = false;
        c->cachedBound = false;
    }
    if (!s.fDiceable && s.type != RH_PRIM_BILINEAR) {
        std::vector<Surf*> s0, s1;
        surf_split(*a, s0);
        surf_split(*b, s1);
        delete a;
        delete b;
        out.insert(out.end(), s0.begin(), s0.end());
        out.insert(out.end(), s1.begin(), s1.end());
        return;
    }
    out.push_back(a);
    out.push_back(b);
}

// CqDeformingSurface::Split (surface.h:657-692): every key splits the same way
// (their m_fDiceable is never cleared, so never twice) and the halves are
// paired into new deforming surfaces (m_fDiceable true, m_SplitCount + 1).
static void surf_split_motion(const Surf& s, std::vector<Surf*>& out) {
    Surf k0 = s;
    k0.motion = false;
    k0.fDiceable = true;
    Surf k1 = key1_of(s);
    std::vector<Surf*> h0, h1;
    surf_split(k0, h0);
    surf_split(k1, h1);
    for (size_t i = 0; i < h0.size(); ++i) {
        Surf* c = h0[i];
        c->motion = true;
        c->keyTime[0] = s.keyTime[0];
        c->keyTime[1] = s.keyTime[1];
        for (int k = 0; k < 16; ++k) c->cp1[k] = h1[i]->cp[k];
        c->matTx1 = h1[i]->matTx;
        c->fDiceable = true;
        c->splitCount = s.splitCount + 1;
        delete h1[i];
        out.push_back(c);
    }
}

// CqForwardDiffBezier (forwarddiff.h:35-103) over V3.
struct FwdDiff {
    float M00, M01, M02, M03, M10, M11, M12, M13, M20, M21, M22, M23;
    V3 f, df, ddf, dddf;
    explicit FwdDiff(float dt) {
        float dt2 = dt * dt, dt3 = dt2 * dt, dt3_3 = 3 * dt3, dt3_6 = 6 * dt3, dt3_18 = 18 * dt3,
              dt2_3 = 3 * dt2, dt2_6 = 6 * dt2, dt_3 = 3 * dt;
        M00 = -dt3_6; M01 = dt3_18; M02 = -dt3_18; M03 = dt3_6;
        M10 = dt2_6 - dt3_6; M11 = dt3_18 - 2 * dt2_6; M12 = dt2_6 - dt3_18; M13 = dt3_6;
        M20 = dt2_3 - dt_3 - dt3; M21 = dt3_3 - dt2_6 + dt_3; M22 = dt2_3 - dt3_3; M23 = dt3;
    }
    void calc(V3 A, V3 B, V3 C, V3 D) {
        f = A;
        df = ((A * M20 + B * M21) + C * M22) + D * M23;
        ddf = ((A * M10 + B * M11) + C * M12) + D * M13;
        dddf = ((A * M00 + B * M01) + C * M02) + D * M03;
    }
    V3 get() {
        V3 res = f;
        f = f + df;
        df = df + ddf;
        ddf = ddf + dddf;
        return res;
    }
};

// CqSurface::Dice -> camera-space grid points.
//  sphere:   DiceAll (quadrics.cpp:82-254): P = m_matTx * (r * unitP)
//  bicubic:  bicubicPatchNatDice (patch.cpp:205-243)
//  bilinear: surfaceNaturalDice with BilinearEvaluate (surface.cpp:365-388, bilinear.h:189-225)
static void surf_dice(const Surf& s, Grid& g) {
    const int uS = s.uDice, vS = s.vDice;
    g.cu = uS; g.cv = vS;
    g.P.resize((size_t)(uS + 1) * (vS + 1));
    if (s.type == RH_PRIM_SPHERE) {
        sphere_dice_points(s, uS, vS, g.P);
        for (auto& p : g.P) p = mat_vec3(s.matTx, p);
    } else if (s.type == RH_PRIM_BICUBIC) {
        float fu = (float)uS, fv = (float)vS;
        FwdDiff vFD0(1.0f / fv), vFD1(1.0f / fv), vFD2(1.0f / fv), vFD3(1.0f / fv), uFD0(1.0f / fu);
        vFD0.calc(s.cp[0], s.cp[4], s.cp[8], s.cp[12]);
        vFD1.calc(s.cp[1], s.cp[5], s.cp[9], s.cp[13]);
        vFD2.calc(s.cp[2], s.cp[6], s.cp[10], s.cp[14]);
        vFD3.calc(s.cp[3], s.cp[7], s.cp[11], s.cp[15]);
        for (int iv = 0; iv <= vS; ++iv) {
            V3 vA = vFD0.get(), vB = vFD1.get(), vC = vFD2.get(), vD = vFD3.get();
            uFD0.calc(vA, vB, vC, vD);
            for (int iu = 0; iu <= uS; ++iu) g.P[(size_t)iv * (uS + 1) + iu] = uFD0.get();
        }
    } else {
        for (int iv = 0; iv <= vS; ++iv) {
            float v = (1.0f / (float)vS) * iv;
            for (int iu = 0; iu <= uS; ++iu) {
                float u = (1.0f / (float)uS) * iu;
                const V3 &A = s.cp[0], &B = s.cp[1], &C = s.cp[2], &D = s.cp[3];
                V3 AB, CD, R;
                if (u <= 0.0) { AB = A; CD = C; }
                else if (u >= 1.0) { AB = B; CD = D; }
                else { AB = (B - A) * u + A; CD = (D - C) * u + C; }
                if (v <= 0.0) R = AB;
                else if (v >= 1.0) R = CD;
                else R = (CD - AB) * v + AB;
                g.P[(size_t)iv * (uS + 1) + iu] = R;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Samples and pixels: SqSampleData / SqImageSample / CqImagePixel
// (imagepixel.h:76-327).
enum { Flag_Matte = 0x0001, Flag_MatteAlpha = 0x0002, Flag_Valid = 0x0004 };
struct Hit {
    float d[7];     // Ci.rgb, Oi.rgb, depth (Sample_Red..Sample_Depth)
    uint32_t id;    // micropolygon id
    uint32_t flags;
};
struct Sample {
    V2 pos;
    V2 dofOffset;   // lens position in the unit disk (setSamples, imagepixel.cpp:357)
    float time, lod;
    float occlZ;
    int occlIndex;
    Hit occl;
    std::vector<Hit> hits;
};

// ---------------------------------------------------------------------------
// Occlusion tree (occlusion.cpp:54-281)
struct OcclTree {
    V2 bmin, bmax;
    std::vector<float> depth;
    int firstLeaf = 0, numLevels = 0;
    bool splitXFirst = true;
    bool needsUpdate = false;

    static int index_for_point(int numLevels, bool splitXFirst, int x, int y) {
        if (!splitXFirst ^ (numLevels % 2 == 1)) std::swap(x, y);
        int index = 1 << (numLevels - 1);
        int idxBit = 0;
        while (x != 0 || y != 0) {
            index |= (x & 1) << idxBit | (y & 1) << (idxBit + 1);
            x >>= 1; y >>= 1; idxBit += 2;
        }
        return index - 1;
    }
    void propagate() {
        for (int i = static_cast<int>(std::pow(2.0, numLevels - 1)) - 2; i >= 0; --i)
            depth[i] = tmax(depth[2 * i + 1], depth[2 * i + 2]);
        needsUpdate = false;
    }
    void update() { if (needsUpdate) propagate(); }
    void set_depth(float d, int index) { depth[index] = d; needsUpdate = true; }
    bool can_cull(const Bound& b) const {
        float tminX = tmax(b.mn.x, bmin.x), tminY = tmax(b.mn.y, bmin.y);
        float tmaxX = tmin(b.mx.x, bmax.x), tmaxY = tmin(b.mx.y, bmax.y);
        struct Node { float minX, minY, maxX, maxY; int index; bool splitInX; };
        std::vector<Node> stack;
        stack.push_back({bmin.x, bmin.y, bmax.x, bmax.y, 0, splitXFirst});
        while (!stack.empty()) {
            Node node = stack.back();
            stack.pop_back();
            if (!(tminX > node.maxX || tminY > node.maxY || tmaxX < node.minX || tmaxY < node.minY)) {
                if (depth[node.index] < b.mn.z) continue;
                if (node.index >= firstLeaf) return false;
                if (node.splitInX) {
                    float avgX = (float)(0.5 * (node.maxX + node.minX));
                    stack.push_back({node.minX, node.minY, avgX, node.maxY, node.index * 2 + 1, !node.splitInX});
                    stack.push_back({avgX, node.minY, node.maxX, node.maxY, node.index * 2 + 2, !node.splitInX});
                } else {
                    float avgY = (float)(0.5 * (node.maxY + node.minY));
                    stack.push_back({node.minX, node.minY, node.maxX, avgY, node.index * 2 + 1, !node.splitInX});
                    stack.push_back({node.minX, avgY, node.maxX, node.maxY, node.index * 2 + 2, !node.splitInX});
                }
            }
        }
        return true;
    }
};

// setupTree's geometry for a sample region of w x h pixels at (sMinX, sMinY)
// with xS x yS samples per pixel: every sampled leaf at FLT_MAX, the padding
// leaves at 0, depths propagated (occlusion.cpp:54-107).
static void init_tree(OcclTree& tree, int sMinX, int sMinY, int w, int h, int xS, int yS) {
    int numX = w * xS, numY = h * yS;
    int depthX = (int)lceil(aq_log2((float)numX));
    int depthY = (int)lceil(aq_log2((float)numY));
    if (depthX < depthY) depthX = depthY - 1;
    else if (depthY < depthX) depthY = depthX - 1;
    tree.splitXFirst = depthX >= depthY;
    int depth = depthX + depthY;
    tree.numLevels = depth + 1;
    int numLeaf = 1 << depth;
    tree.firstLeaf = numLeaf - 1;
    tree.depth.assign(2 * numLeaf - 1, 0.0f);
    tree.bmin = v2((float)sMinX, (float)sMinY);
    V2 diag = v2((float)w, (float)h);
    V2 scale = v2(float(1 << depthX) / numX, float(1 << depthY) / numY);
    tree.bmax = tree.bmin + v2(diag.x * scale.x, diag.y * scale.y);
    for (int y = 0; y < numY; ++y)
        for (int x = 0; x < numX; ++x)
            tree.depth[OcclTree::index_for_point(tree.numLevels, tree.splitXFirst, x, y)] = FLT_MAX;
    tree.propagate();
}

// ---------------------------------------------------------------------------
struct Bucket {
    int col = 0, row = 0, xPos = 0, yPos = 0, xSize = 0, ySize = 0;
    bool processed = false;
    std::vector<Surf*> heap;
    std::vector<MP*> mps;
};

// closest_surface (bucket.h:164-176)
struct ClosestSurface {
    bool operator()(const Surf* s1, const Surf* s2) const {
        if (s1->cachedBound && s2->cachedBound) return s1->rasterBound.mn.z > s2->rasterBound.mn.z;
        return true;
    }
};

struct Hider {
    Frame F;
    const rh_prim* prims = nullptr;
    bool cull = true;
    Rng rng;
    SamplerTables samp;
    std::vector<Bucket> buckets;  // row-major, cYBuckets * cXBuckets
    std::vector<Sample> samples;  // extended image, extW*extH*nsamp
    std::vector<int> dofIndex;    // per pixel: m_DofOffsetIndices (lens cell -> sample)
    std::vector<uint8_t> pixelHasValid;
    std::vector<float> channels;  // display region, 9 per pixel
    std::vector<uint8_t> rgba8;   // display region
    std::vector<float> ditherRec; // display 0's dither draw per display pixel
    int ditherIndex = 0;          // the display whose draws ditherRec records
    uint32_t mpCounter = 0;
    rh_stats st;
    // current bucket state
    int sMinX = 0, sMinY = 0, sMaxX = 0, sMaxY = 0;  // sample region
    OcclTree tree;

    Bucket& bucket(int c, int r) { return buckets[(size_t)r * F.cXBuckets + c]; }
    Sample& sample_at(int px, int py, int i) {
        return samples[((size_t)(py - F.extY0) * F.extW + (px - F.extX0)) * F.nsamp + i];
    }

    void set_image() {
        const rh_options& o = F.o;
        buckets.assign((size_t)F.cXBuckets * F.cYBuckets, Bucket());
        int rowPos = 0;
        for (int r = 0; r < F.cYBuckets; ++r) {
            int rowSize = tmin(o.yres - rowPos, o.bucket_y);
            int colPos = 0;
            for (int c = 0; c < F.cXBuckets; ++c) {
                Bucket& b = bucket(c, r);
                b.col = c; b.row = r;
                b.xPos = colPos; b.yPos = rowPos;
                b.xSize = tmin(o.xres - colPos, o.bucket_x);
                b.ySize = rowSize;
                colPos += o.bucket_x;
            }
            rowPos += o.bucket_y;
        }
    }

    // CqImageBuffer::CullSurface (imagebuffer.cpp:268-345)
    bool cull_surface(Bound B, Surf* s) {
        const rh_options& o = F.o;
        if (B.mn.z >= o.clip_far || B.mx.z <= o.clip_near) return true;
        if (B.mn.z <= FLT_EPSILON) {
            s->fDiceable = false;
            if (s->splitCount > o.eyesplits) return true;
            return false;
        }
        float minz = B.mn.z, maxz = B.mx.z;
        B = bound_transform(B, F.camToRaster);
        if (F.dof.on) {
            // grown by the larger circle of confusion at its depths (imagebuffer.cpp:312-322)
            const V2 minZCoc = circle_of_confusion(F.dof, minz), maxZCoc = circle_of_confusion(F.dof, maxz);
            const float cocX = tmax(minZCoc.x, maxZCoc.x), cocY = tmax(minZCoc.y, maxZCoc.y);
            B.mn.x = B.mn.x - cocX;
            B.mn.y = B.mn.y - cocY;
            B.mx.x = B.mx.x + cocX;
            B.mx.y = B.mx.y + cocY;
        }
        B.mn.x = B.mn.x - o.filter_xwidth / 2.0f;
        B.mn.y = B.mn.y - o.filter_ywidth / 2.0f;
        B.mx.x = B.mx.x + o.filter_xwidth / 2.0f;
        B.mx.y = B.mx.y + o.filter_ywidth / 2.0f;
        if (B.mn.x > F.cropXMax || B.mn.y > F.cropYMax || B.mx.x < F.cropXMin || B.mx.y < F.cropYMin)
            return true;
        B.mn.z = minz;
        B.mx.z = maxz;
        s->rasterBound = B;
        s->cachedBound = true;
        return false;
    }

    void add_gprim(Bucket& b, Surf* s) {
        b.heap.push_back(s);
        std::push_heap(b.heap.begin(), b.heap.end(), ClosestSurface());
    }

    // CqImageBuffer::PostSurface (imagebuffer.cpp:353-446)
    void post_surface(Surf* s) {
        st.surfaces++;
        Bound B = surf_bound(*s);
        if (cull_surface(B, s)) { delete s; return; }
        B = s->rasterBound;
        int XMinb = 0, YMinb = 0, XMaxb = 0, YMaxb = 0;
        if (s->fDiceable) {
            XMinb = static_cast<int>(B.mn.x) / F.o.bucket_x;
            YMinb = static_cast<int>(B.mn.y) / F.o.bucket_y;
            XMaxb = static_cast<int>(B.mx.x) / F.o.bucket_x;
            YMaxb = static_cast<int>(B.mx.y) / F.o.bucket_y;
        }
        XMinb = tclamp(XMinb, F.bxMin, F.bxMax - 1);
        YMinb = tclamp(YMinb, F.byMin, F.byMax - 1);
        XMaxb = tclamp(XMaxb, F.bxMin, F.bxMax - 1);
        YMaxb = tclamp(YMaxb, F.byMin, F.byMax - 1);
        Bucket* bk = &bucket(XMinb, YMinb);
        if (bk->processed) {
            int yb = YMinb, xb = XMinb + 1;
            bool done = false;
            while (!done && yb <= YMaxb) {
                while (!done && xb <= XMaxb) {
                    Bucket& avail = bucket(xb, yb);
                    if (!avail.processed) { add_gprim(avail, s); done = true; }
                    ++xb;
                }
                xb = XMinb;
                ++yb;
            }
            if (!done) delete s;
        } else {
            add_gprim(*bk, s);
        }
    }

    // CqImageBuffer::RepostSurface (imagebuffer.cpp:449-507)
    void repost_surface(const Bucket& old, Surf* s) {
        const Bound& rb = s->rasterBound;
        int nextX = old.col + 1, nextY = old.row;
        int xpos = old.xPos + old.xSize;
        if (nextX < F.bxMax && rb.mx.x >= xpos) { add_gprim(bucket(nextX, nextY), s); return; }
        ++nextY;
        nextX = (int)tmax<long>(F.bxMin, lfloor(rb.mn.x) / F.o.bucket_x);
        int ypos = old.yPos + old.ySize;
        if (nextX < F.bxMax && nextY < F.byMax && rb.mx.y >= ypos) { add_gprim(bucket(nextX, nextY), s); return; }
        delete s;
    }

    // CqImageBuffer::AddMPG (imagebuffer.cpp:514-591)
    void add_mpg(MP* mp) {
        const rh_options& o = F.o;
        Bound B = mp->bound;
        if (F.dof.on) {
            // the larger circle of confusion at its two depths (imagebuffer.cpp:519-530)
            const V2 c0 = circle_of_confusion(F.dof, B.mn.z), c1 = circle_of_confusion(F.dof, B.mx.z);
            const V2 maxCoC = v2(tmax(c0.x, c1.x), tmax(c0.y, c1.y));
            B.mn.x = B.mn.x - maxCoC.x; B.mn.y = B.mn.y - maxCoC.y; B.mn.z = B.mn.z - 0.0f;
            B.mx.x = B.mx.x + maxCoC.x; B.mx.y = B.mx.y + maxCoC.y; B.mx.z = B.mx.z + 0.0f;
        }
        if (B.mx.x < F.cropXMin - o.filter_xwidth / 2.0f || B.mx.y < F.cropYMin - o.filter_ywidth / 2.0f ||
            B.mn.x > F.cropXMax + o.filter_xwidth / 2.0f || B.mn.y > F.cropYMax + o.filter_ywidth / 2.0f)
            return;
        B.mn.x = B.mn.x - (float)lfloor(o.filter_xwidth / 2.0f);
        B.mn.y = B.mn.y - (float)lfloor(o.filter_ywidth / 2.0f);
        B.mx.x = B.mx.x + (float)lfloor(o.filter_xwidth / 2.0f);
        B.mx.y = B.mx.y + (float)lfloor(o.filter_ywidth / 2.0f);
        int iXBa = static_cast<int>(B.mn.x / o.bucket_x);
        int iYBa = static_cast<int>(B.mn.y / o.bucket_y);
        int iXBb = static_cast<int>(B.mx.x / o.bucket_x);
        int iYBb = static_cast<int>(B.mx.y / o.bucket_y);
        if (iXBb < F.bxMin || iYBb < F.byMin || iXBa >= F.bxMax || iYBa >= F.byMax) return;
        if (iXBa < F.bxMin) iXBa = F.bxMin;
        if (iYBa < F.byMin) iYBa = F.byMin;
        if (iXBb >= F.bxMax) iXBb = F.bxMax - 1;
        if (iYBb >= F.byMax) iYBb = F.byMax - 1;
        for (int i = iXBa; i <= iXBb; ++i)
            for (int j = iYBa; j <= iYBb; ++j) {
                Bucket& b = bucket(i, j);
                if (!b.processed) {
                    b.mps.push_back(mp);
                    mp->refs++;
                    st.mp_bucket_refs++;
                }
            }
    }

    // CqMicroPolygon::ComputeVertexOrder (micropolygon.cpp:1207-1284)
    static int vertex_order(const Grid& g, int index) {
        int cu = g.cu;
        int IndexA = index, IndexB = index + 1, IndexC = index + cu + 2, IndexD = index + cu + 1;
        int CodeA = 0, CodeB = 1, CodeC = 3, CodeD = 2;
        const V3* pP = g.P.data();
        if ((double)mag2(pP[IndexA] - pP[IndexB]) < 1e-8) {
            IndexB = IndexC; CodeB = CodeC; IndexC = IndexD; CodeC = CodeD; IndexD = -1; CodeD = -1;
        } else if ((double)mag2(pP[IndexB] - pP[IndexC]) < 1e-8) {
            IndexB = IndexC; CodeB = CodeC; IndexC = IndexD; CodeC = CodeD; IndexD = -1; CodeD = -1;
        } else if ((double)mag2(pP[IndexC] - pP[IndexD]) < 1e-8) {
            IndexC = IndexD; CodeC = CodeD; IndexD = -1; CodeD = -1;
        } else if ((double)mag2(pP[IndexD] - pP[IndexA]) < 1e-8) {
            IndexD = IndexC; CodeD = CodeC; IndexD = -1; CodeD = -1;
        }
        const V3& vA2 = pP[IndexA];
        const V3& vB2 = pP[IndexB];
        const V3& vC2 = pP[IndexC];
        bool fFlip = ((vA2.x - vB2.x) * (vB2.y - vC2.y)) >= ((vA2.y - vB2.y) * (vB2.x - vC2.x));
        int code;
        if (!fFlip)
            code = (CodeD == -1) ? ((CodeA & 0x3) | ((CodeC & 0x3) << 2) | ((CodeB & 0x3) << 4) | Degeneracy_Mask)
                                 : ((CodeA & 0x3) | ((CodeD & 0x3) << 2) | ((CodeC & 0x3) << 4) | ((CodeB & 0x3) << 6));
        else
            code = (CodeD == -1) ? ((CodeA & 0x3) | ((CodeB & 0x3) << 2) | ((CodeC & 0x3) << 4) | Degeneracy_Mask)
                                 : ((CodeA & 0x3) | ((CodeB & 0x3) << 2) | ((CodeC & 0x3) << 4) | ((CodeD & 0x3) << 6));
        return code;
    }

    // RenderSurface: dice -> shade (constant) -> bust (bucketprocessor.cpp:942-1059,
    // micropolygon.cpp:345-527, 641-892).
    void dice_shade_bust(Surf* s) {
        if (s->motion) {
            dice_shade_bust_motion(s);
            return;
        }
        const rh_prim& p = prims[s->prim];
        Grid* g = new Grid();
        surf_dice(*s, *g);
        st.grids++;
        // constant shader: Oi = Os; Ci = Os * Cs (shaders/surface/constant.sl)
        for (int k = 0; k < 3; ++k) { g->Oi[k] = p.Os[k]; g->Ci[k] = p.Os[k] * p.Cs[k]; }
        g->constantShading = p.shading_interpolation == RH_SHADING_CONSTANT;
        g->matteFlag = p.matte == 0 ? 0u : (p.matte == 2 ? (uint32_t)Flag_MatteAlpha : (uint32_t)Flag_Matte);
        // Transparency culling (micropolygon.cpp:497-520): Oi == black culls every MP.
        bool allBlack = (g->Oi[0] == 0.0f && g->Oi[1] == 0.0f && g->Oi[2] == 0.0f);
        int cu = g->cu, cv = g->cv;
        if (allBlack) { delete g; return; }
        // Projection to hybrid raster space, keeping camera z (micropolygon.cpp:723-731).
        for (int i = (int)g->P.size() - 1; i >= 0; --i) {
            float zdepth = g->P[i].z;
            V3 r = mat_vec3(F.camToRaster, g->P[i]);
            r.z = zdepth;
            g->P[i] = r;
        }
        g->mpBase = mpCounter;
        g->refs = 1;
        for (int iv = 0; iv < cv; ++iv)
            for (int iu = 0; iu < cu; ++iu) {
                int idx = iv * (cu + 1) + iu;
                MP* mp = new MP();
                mp->g = g;
                g->refs++;
                mp->index = idx;
                mp->id = mpCounter++;
                st.micropolygons++;
                mp->indexCode = vertex_order(*g, idx);
                // CalculateBound (micropolygon.cpp:1663-1673)
                const V3& A = g->P[idx];
                const V3& B = g->P[idx + 1];
                const V3& C = g->P[idx + cu + 2];
                const V3& D = g->P[idx + cu + 1];
                mp->bound.mn = vmin(vmin(vmin(A, B), C), D);
                mp->bound.mx = vmax(vmax(vmax(A, B), C), D);
                mp->refs = 1;
                add_mpg(mp);
                release_mp(mp);
            }
        release_grid(g);
    }
    // CqDeformingSurface::Dice (surface.h:643-655) -> CqMotionMicroPolyGrid:
    // every key diced with key 0's sizes; Shade(false) shades key 0 and never
    // culls the grid (micropolygon.cpp:920-924, 493-520), and its Split
    // busts every quad (no transparency cull) into a CqMicroPolygonMotion
    // with one key per grid key (micropolygon.cpp:946-1156; two-sided
    // surfaces, so no backface cull).
    void dice_shade_bust_motion(Surf* s) {
        const rh_prim& p = prims[s->prim];
        Grid* g = new Grid();
        surf_dice(*s, *g);
        Grid g1;
        surf_dice(key1_of(*s), g1);
        st.grids++;
        for (int k = 0; k < 3; ++k) { g->Oi[k] = p.Os[k]; g->Ci[k] = p.Os[k] * p.Cs[k]; }
        g->constantShading = p.shading_interpolation == RH_SHADING_CONSTANT;
        g->matteFlag = p.matte == 0 ? 0u : (p.matte == 2 ? (uint32_t)Flag_MatteAlpha : (uint32_t)Flag_Matte);
        const int cu = g->cu, cv = g->cv;
        for (Grid* gk : {g, &g1})
            for (int i = (int)gk->P.size() - 1; i >= 0; --i) {
                float zdepth = gk->P[i].z;
                V3 r = mat_vec3(F.camToRaster, gk->P[i]);
                r.z = zdepth;
                gk->P[i] = r;
            }
        g->mpBase = mpCounter;
        g->refs = 1;
        for (int iv = 0; iv < cv; ++iv)
            for (int iu = 0; iu < cu; ++iu) {
                int idx = iv * (cu + 1) + iu;
                MP* mp = new MP();
                mp->g = g;
                g->refs++;
                mp->index = idx;
                mp->id = mpCounter++;
                st.micropolygons++;
                mp->motion = true;
                const Grid* keys[2] = {g, &g1};
                for (int k = 0; k < 2; ++k) {
                    const std::vector<V3>& P = keys[k]->P;
                    mp->K[k][0] = P[idx];
                    mp->K[k][1] = P[idx + 1];
                    mp->K[k][2] = P[idx + cu + 1];
                    mp->K[k][3] = P[idx + cu + 2];
                    mp->T[k] = s->keyTime[k];
                    // CqMovingMicroPolygonKey::GetBound (micropolygon.cpp:1974-1989)
                    const V3* q = mp->K[k];
                    Bound& B = mp->keyBound[k];
                    B.mn = v3(tmin(q[0].x, tmin(q[1].x, tmin(q[2].x, q[3].x))),
                              tmin(q[0].y, tmin(q[1].y, tmin(q[2].y, q[3].y))),
                              tmin(q[0].z, tmin(q[1].z, tmin(q[2].z, q[3].z))));
                    B.mx = v3(tmax(q[0].x, tmax(q[1].x, tmax(q[2].x, q[3].x))),
                              tmax(q[0].y, tmax(q[1].y, tmax(q[2].y, q[3].y))),
                              tmax(q[0].z, tmax(q[1].z, tmax(q[2].z, q[3].z))));
                }
                // AppendKey (micropolygon.cpp:1952-1967): m_Bound is key 0's
                // bound encapsulating key 1's
                mp->bound = mp->keyBound[0];
                const Bound& B1 = mp->keyBound[1];
                mp->bound.mx = v3(tmax(mp->bound.mx.x, B1.mx.x), tmax(mp->bound.mx.y, B1.mx.y),
                                  tmax(mp->bound.mx.z, B1.mx.z));
                mp->bound.mn = v3(tmin(mp->bound.mn.x, B1.mn.x), tmin(mp->bound.mn.y, B1.mn.y),
                                  tmin(mp->bound.mn.z, B1.mn.z));
                // Initialise: the vertex order of the primary key (key 0)
                mp->indexCode = vertex_order(*g, idx);
                mp->refs = 1;
                add_mpg(mp);
                release_mp(mp);
            }
        release_grid(g);
    }

    void release_grid(Grid* g) { if (--g->refs == 0) delete g; }
    void release_mp(MP* mp) {
        if (--mp->refs == 0) { release_grid(mp->g); delete mp; }
    }

    // --- Hit testing -------------------------------------------------------
    struct HitCache {
        V3 P[4];
        float z[4];
        float YM[4], XM[4], X[4], Y[4];
        int lastFailed;
        V2 A, E, Fv, G;
        bool linear;
    };
    // cachePointInPolyTest (micropolygon.cpp:1346-1392) and
    // CqInvBilinear::setVertices (bilinear.h:260-275)
    static void cache_point_in_poly(HitCache& c, const V3* pin, int code) {
        for (int i = 0; i < 4; ++i) c.z[i] = pin[i].z;
        V2 A = v2(pin[0].x, pin[0].y), B = v2(pin[1].x, pin[1].y), C = v2(pin[2].x, pin[2].y),
           D = v2(pin[3].x, pin[3].y);
        c.A = A;
        c.E = B - A;
        c.Fv = C - A;
        V2 negE = v2(-c.E.x, -c.E.y);
        c.G = (negE - C) + D;
        c.linear = false;
        float patchSize = tmax(maxNorm(c.Fv), maxNorm(c.E));
        float irregularity = maxNorm(c.G);
        if ((double)irregularity < 1e-2 * (double)patchSize) c.linear = true;
        V3 pts[4] = {pin[(code >> 2) & 0x3], pin[(code >> 4) & 0x3], pin[(code >> 6) & 0x3], pin[code & 0x3]};
        int j = 3;
        for (int i = 0; i < 4; ++i) {
            c.YM[i] = pts[i].x - pts[j].x;
            c.XM[i] = pts[i].y - pts[j].y;
            c.X[i] = pts[j].x;
            c.Y[i] = pts[j].y;
            j = i;
        }
        if (code & Degeneracy_Mask) {
            for (int i = 2; i < 4; ++i) {
                c.YM[i] = pts[3].x - pts[1].x;
                c.XM[i] = pts[3].y - pts[1].y;
                c.X[i] = pts[1].x;
                c.Y[i] = pts[1].y;
            }
        }
        c.lastFailed = 0;
    }
    // CqInvBilinear::solve (bilinear.h:298-305)
    static V2 solve(bool unsafeInvert, V2 M1, V2 M2, V2 b) {
        float det = cross2(M1, M2);
        if (unsafeInvert || det != 0) det = 1 / det;
        return det * v2(cross2(b, M2), -cross2(b, M1));
    }
    // bilinEval (bilinear.h:307-310)
    static V2 bilin_eval(const HitCache& c, V2 uv) {
        return ((c.A + c.E * uv.x) + c.Fv * uv.y) + (c.G * uv.x) * uv.y;
    }
    // CqInvBilinear::operator() (bilinear.h:277-290)
    static V2 inv_bilinear(const HitCache& c, V2 P) {
        V2 uv = v2(0.5f, 0.5f);
        uv = uv - solve(true, c.E + c.G * uv.y, c.Fv + c.G * uv.x, bilin_eval(c, uv) - P);
        if (!c.linear) uv = uv - solve(false, c.E + c.G * uv.y, c.Fv + c.G * uv.x, bilin_eval(c, uv) - P);
        return uv;
    }
    // bilerp (bilinear.h:228-236)
    static float bilerp(float A, float B, float C, float D, V2 uv) {
        float w0 = (1 - uv.y) * (1 - uv.x);
        float w1 = (1 - uv.y) * uv.x;
        float w2 = uv.y * (1 - uv.x);
        float w3 = uv.y * uv.x;
        return ((w0 * A + w1 * B) + w2 * C) + w3 * D;
    }
    // fContains (micropolygon.cpp:1295-1342)
    static bool contains(HitCache& c, V2 p, float& depth, V2& uv) {
        float x = p.x, y = p.y;
        int e = c.lastFailed;
        for (int i = 0; i < 4; ++i) {
            float t = ((y - c.Y[e]) * c.YM[e]) - ((x - c.X[e]) * c.XM[e]);
            if (e & 2) {
                if (t < 0) { c.lastFailed = e; return false; }
            } else {
                if (t <= 0) { c.lastFailed = e; return false; }
            }
            e = (e + 1) & 3;
        }
        uv = inv_bilinear(c, p);
        depth = bilerp(c.z[0], c.z[1], c.z[2], c.z[3], uv);
        return true;
    }

    // RenderMPG_Static (bucketprocessor.cpp:1097-1218)
    void render_mp_static(MP* mp) {
        const Grid& g = *mp->g;
        HitCache hc;
        int cu = g.cu;
        hc.P[0] = g.P[mp->index];
        hc.P[1] = g.P[mp->index + 1];
        hc.P[2] = g.P[mp->index + cu + 1];
        hc.P[3] = g.P[mp->index + cu + 2];
        cache_point_in_poly(hc, hc.P, mp->indexCode);
        // CacheOutputInterpCoeffsSmooth: four equal vertex colours
        // (micropolygon.cpp:1494-1529); isOpaque = Oi >= (1,1,1).
        bool isOpaque = g.Oi[0] >= 1.0f && g.Oi[1] >= 1.0f && g.Oi[2] >= 1.0f;
        const Bound& Bd = mp->bound;
        float bminx = Bd.mn.x, bmaxx = Bd.mx.x, bminy = Bd.mn.y, bmaxy = Bd.mx.y;
        int eX = (int)lceil(bmaxx), eY = (int)lceil(bmaxy);
        if (eX > sMaxX) eX = sMaxX;
        if (eY > sMaxY) eY = sMaxY;
        int sX = static_cast<int>(std::floor(bminx)), sY = static_cast<int>(std::floor(bminy));
        if (sY < sMinY) sY = sMinY;
        if (sX < sMinX) sX = sMinX;
        const int xs = F.nx, ys = F.ny;
        int im = (bminx < sX) ? 0 : static_cast<int>(std::floor((bminx - sX) * xs));
        int in = (bminy < sY) ? 0 : static_cast<int>(std::floor((bminy - sY) * ys));
        int em = (bmaxx > eX) ? xs : (int)lceil((bmaxx - (eX - 1)) * xs);
        int en = (bmaxy > eY) ? ys : (int)lceil((bmaxy - (eY - 1)) * ys);
        if (sX >= eX || sY >= eY) return;
        for (int iY = sY; iY < eY; ++iY)
            for (int iX = sX; iX < eX; ++iX) {
                int n = (iY == sY) ? in : 0;
                int end_n = (iY == (eY - 1)) ? en : ys;
                int start_m = (iX == sX) ? im : 0;
                int end_m = (iX == (eX - 1)) ? em : xs;
                int index_start = n * xs + start_m;
                for (; n < end_n; ++n) {
                    int index = index_start;
                    for (int m = start_m; m < end_m; ++m, ++index) {
                        Sample& sd = sample_at(iX, iY, index);
                        V2 vecP = sd.pos;
                        st.spl_count++;
                        // CqBound::Contains2D (bound.h:144-151)
                        if (vecP.x < Bd.mn.x || vecP.x > Bd.mx.x || vecP.y < Bd.mn.y || vecP.y > Bd.mx.y) continue;
                        if (F.cullableHits && Bd.mn.z > sd.occlZ) continue;
                        // LevelOfDetailBounds default [0,1] (attributes.cpp:125)
                        if (0.0f > sd.lod || sd.lod >= 1.0f) continue;
                        st.spl_bound_hits++;
                        float D;
                        V2 uv;
                        if (contains(hc, vecP, D, uv)) {
                            st.inside++;
                            if (!hc.linear) st.inside_nonlinear++;
                            store_sample(mp, iX, iY, index, D, uv, isOpaque);
                        }
                    }
                    index_start += xs;
                }
            }
    }

    // RenderMPG_MBOrDof for a static micropolygon with depth of field
    // (bucketprocessor.cpp:1221-1469, the UsingDof branches).  Each lens cell
    // shifts the micropolygon's bound by the cell's bound of lens offsets
    // times the larger circle of confusion at its depths; a pixel of the
    // shifted bound offers the one sample whose lens offset lies in that
    // cell (GetDofOffsetIndex).  Every visit counts into SPL_count.
    void render_mp_dof(MP* mp) {
        const Grid& g = *mp->g;
        HitCache hc;
        const int cu = g.cu;
        // CacheHitTestValues with DoF (micropolygon.cpp:1394-1421)
        hc.P[0] = g.P[mp->index];
        hc.P[1] = g.P[mp->index + 1];
        hc.P[2] = g.P[mp->index + cu + 1];
        hc.P[3] = g.P[mp->index + cu + 2];
        V2 cocMult[4];
        for (int i = 0; i < 4; ++i) cocMult[i] = circle_of_confusion(F.dof, hc.P[i].z);
        const V2 cm01 = v2(tmin(cocMult[0].x, cocMult[1].x), tmin(cocMult[0].y, cocMult[1].y));
        const V2 cm23 = v2(tmin(cocMult[2].x, cocMult[3].x), tmin(cocMult[2].y, cocMult[3].y));
        const V2 cocMultMin = v2(tmin(cm01.x, cm23.x), tmin(cm01.y, cm23.y));
        const V2 cx01 = v2(tmax(cocMult[0].x, cocMult[1].x), tmax(cocMult[0].y, cocMult[1].y));
        const V2 cx23 = v2(tmax(cocMult[2].x, cocMult[3].x), tmax(cocMult[2].y, cocMult[3].y));
        const V2 cocMultMax = v2(tmax(cx01.x, cx23.x), tmax(cx01.y, cx23.y));
        const bool isOpaque = g.Oi[0] >= 1.0f && g.Oi[1] >= 1.0f && g.Oi[2] >= 1.0f;
        const Bound& Bd = mp->bound;   // SubBound(0) of a static micropolygon
        const float bminz = Bd.mn.z, bmaxz = Bd.mx.z;
        if (bminz > F.o.clip_far || bmaxz < F.o.clip_near) return;
        const float mpgbminx = Bd.mn.x, mpgbmaxx = Bd.mx.x, mpgbminy = Bd.mn.y, mpgbmaxy = Bd.mx.y;
        const V2 minZCoc = circle_of_confusion(F.dof, bminz), maxZCoc = circle_of_confusion(F.dof, bmaxz);
        const float maxCocX = tmax(minZCoc.x, maxZCoc.x), maxCocY = tmax(minZCoc.y, maxZCoc.y);
        for (int cell = 0; cell < F.nsamp; ++cell) {
            const Bound& DB = F.dofBounds[cell];
            const float leftOffset = DB.mx.x * maxCocX, rightOffset = DB.mn.x * maxCocX;
            const float topOffset = DB.mx.y * maxCocY, bottomOffset = DB.mn.y * maxCocY;
            const float bminx = mpgbminx - leftOffset, bmaxx = mpgbmaxx - rightOffset;
            const float bminy = mpgbminy - topOffset, bmaxy = mpgbmaxy - bottomOffset;
            int eX = (int)lceil(bmaxx), eY = (int)lceil(bmaxy);
            if (eX > sMaxX) eX = sMaxX;
            if (eY > sMaxY) eY = sMaxY;
            int sX = static_cast<int>(std::floor(bminx)), sY = static_cast<int>(std::floor(bminy));
            if (sY < sMinY) sY = sMinY;
            if (sX < sMinX) sX = sMinX;
            if (sX >= eX || sY >= eY) continue;
            for (int iY = sY; iY < eY; ++iY)
                for (int iX = sX; iX < eX; ++iX) {
                    const int index = dofIndex[((size_t)(iY - F.extY0) * F.extW + (iX - F.extX0)) * F.nsamp + cell];
                    Sample& sd = sample_at(iX, iY, index);
                    const V2 vecP = sd.pos;
                    st.spl_count++;
                    // the shifted bound's Contains2D (bound.h:144-151)
                    if (vecP.x < bminx || vecP.x > bmaxx || vecP.y < bminy || vecP.y > bmaxy) continue;
                    if (F.cullableHits && Bd.mn.z > sd.occlZ) continue;
                    if (0.0f > sd.lod || sd.lod >= 1.0f) continue;
                    st.spl_bound_hits++;
                    // CqMicroPolygon::Sample with DoF (micropolygon.cpp:1561-1593):
                    // dofSampleInBound (:1531-1550) on the micropolygon's bound
                    const V2 o = sd.dofOffset;
                    V2 cocMin = v2(vecP.x + cocMultMin.x * o.x, vecP.y + cocMultMin.y * o.y);
                    V2 cocMax = v2(vecP.x + cocMultMax.x * o.x, vecP.y + cocMultMax.y * o.y);
                    if (o.x < 0) std::swap(cocMin.x, cocMax.x);
                    if (o.y < 0) std::swap(cocMin.y, cocMax.y);
                    if (cocMin.x > Bd.mx.x || cocMin.y > Bd.mx.y || cocMax.x < Bd.mn.x || cocMax.y < Bd.mn.y) continue;
                    // the vertices moved against the lens offset, then the
                    // point-in-polygon set-up on them
                    V3 pts[4];
                    for (int i = 0; i < 4; ++i)
                        pts[i] = v3(hc.P[i].x - cocMult[i].x * o.x, hc.P[i].y - cocMult[i].y * o.y, hc.P[i].z - 0.0f);
                    cache_point_in_poly(hc, pts, mp->indexCode);
                    float D;
                    V2 uv;
                    if (contains(hc, vecP, D, uv)) {
                        st.inside++;
                        if (!hc.linear) st.inside_nonlinear++;
                        store_sample(mp, iX, iY, index, D, uv, isOpaque);
                    }
                }
        }
    }

    // CqMicroPolygonMotion::BuildBoundList (micropolygon.cpp:1689-1756) for
    // two keys: the number of sub-bounds is the micropolygon's raster move in
    // its own widths, at most timeRanges; each sub-bound encapsulates the
    // bounds interpolated at its two end times (extrapolated past the last
    // key), and carries its start time.
    void build_bound_list(MP* mp, int timeRanges) {
        const float opentime = F.o.shutter_open, closetime = F.o.shutter_close;
        const Bound& kb0 = mp->keyBound[0];
        const Bound& kb1 = mp->keyBound[1];
        const float polyLen2 = mag2_2d(kb0.mx.x - kb0.mn.x, kb0.mx.y - kb0.mn.y);
        const float moveDist2 = mag2_2d(mp->K[0][0].x - mp->K[1][0].x, mp->K[0][0].y - mp->K[1][0].y);
        const float lengths = std::sqrt(moveDist2 / polyLen2);
        // lfloor of an infinite or NaN ratio is LONG_MIN on x86-64 (the
        // conversion's "integer indefinite"), so max(1, .) gives 1
        long moved = (lengths < 9.2e18f) ? lfloor(lengths) : LONG_MIN;
        if (moved < 1) moved = 1;
        const int divisions = (int)tmin<long>(moved, timeRanges);
        const float dt = (closetime - opentime) / divisions;
        float time = opentime + dt;
        Bound bound = kb0;
        mp->subBound.resize(divisions);
        mp->subTime.resize(divisions);
        for (int i = 0; i < divisions; ++i) {
            const float mix = (time - mp->T[0]) / (mp->T[1] - mp->T[0]);
            Bound mid = kb0;
            mid.mn = mid.mn + mix * (kb1.mn - kb0.mn);
            mid.mx = mid.mx + mix * (kb1.mx - kb0.mx);
            bound.mx = v3(tmax(bound.mx.x, mid.mx.x), tmax(bound.mx.y, mid.mx.y), tmax(bound.mx.z, mid.mx.z));
            bound.mn = v3(tmin(bound.mn.x, mid.mn.x), tmin(bound.mn.y, mid.mn.y), tmin(bound.mn.z, mid.mn.z));
            mp->subBound[i] = bound;
            mp->subTime[i] = time - dt;
            bound = mid;
            time += dt;
        }
    }

    // CqMicroPolygonMotion::Sample (micropolygon.cpp:1768-1914) without DoF
    // or triangle split lines: the tight bound and the vertices at the
    // sample's time, then the static point-in-polygon test on them.
    bool sample_motion(const MP* mp, V2 vecSample, float time, float& D, V2& uv, bool& linear) {
        int iIndex = 0;
        float Fraction = 0.0f;
        bool Exact = true;
        if (time > mp->T[0]) {
            if (time >= mp->T[1]) {
                iIndex = 1;
            } else {
                Fraction = (time - mp->T[0]) / (mp->T[1] - mp->T[0]);
                Exact = (mp->T[0] == time);
            }
        }
        Bound tight;
        if (Exact) {
            tight = mp->keyBound[iIndex];
        } else {
            const Bound& b1 = mp->keyBound[0];
            const Bound& b2 = mp->keyBound[1];
            tight.mn = (1 - Fraction) * b1.mn + Fraction * b2.mn;
            tight.mx = (1 - Fraction) * b1.mx + Fraction * b2.mx;
        }
        if (vecSample.x < tight.mn.x || vecSample.x > tight.mx.x || vecSample.y < tight.mn.y ||
            vecSample.y > tight.mx.y)
            return false;
        V3 pts[4];
        if (Exact) {
            for (int i = 0; i < 4; ++i) pts[i] = mp->K[iIndex][i];
        } else {
            const float F1 = 1.0f - Fraction;
            for (int i = 0; i < 4; ++i) pts[i] = (F1 * mp->K[0][i]) + (Fraction * mp->K[1][i]);
        }
        HitCache hc;
        cache_point_in_poly(hc, pts, mp->indexCode);
        linear = hc.linear;
        return contains(hc, vecSample, D, uv);
    }

    // RenderMPG_MBOrDof for a moving micropolygon without DoF
    // (bucketprocessor.cpp:1221-1469).  A sub-bound visits the samples of its
    // pixels whose index lies in the window its times give (sample i's time
    // lies in the i-th of n equal parts of the shutter); every visit counts
    // into SPL_count, and samples outside the sub-bound's times are skipped.
    // An index window reaching past a pixel's samples (the reference reads
    // beyond them) stops at the last sample.
    void render_mp_motion(MP* mp) {
        const Grid& g = *mp->g;
        const bool isOpaque = g.Oi[0] >= 1.0f && g.Oi[1] >= 1.0f && g.Oi[2] >= 1.0f;
        const float opentime = F.o.shutter_open, closetime = F.o.shutter_close;
        const int numSamples = F.nsamp;
        const bool fastShutter = isClose(closetime, opentime);
        const float timePerSample = fastShutter ? 0.0f : numSamples / (closetime - opentime);
        const int timeRanges = tmax(4, F.nx * F.ny);
        if (mp->subBound.empty()) build_bound_list(mp, timeRanges);
        const int nb = (int)mp->subBound.size();
        for (int b = 0; b < nb; ++b) {
            const Bound& Bd = mp->subBound[b];
            const float time0 = mp->subTime[b];
            const float time1 = (b != nb - 1) ? mp->subTime[b + 1] : closetime;
            if (time1 < opentime || time0 > closetime) continue;
            int indexT0, indexT1;
            if (fastShutter) {
                indexT0 = 0;
                indexT1 = numSamples;
            } else {
                indexT0 = (int)tmax<long>(0, lfloor((time0 - opentime) * timePerSample));
                indexT1 = (int)lceil((time1 - opentime) * timePerSample);
            }
            if (Bd.mn.z > F.o.clip_far || Bd.mx.z < F.o.clip_near) continue;
            int eX = (int)lceil(Bd.mx.x), eY = (int)lceil(Bd.mx.y);
            if (eX > sMaxX) eX = sMaxX;
            if (eY > sMaxY) eY = sMaxY;
            int sX = static_cast<int>(std::floor(Bd.mn.x)), sY = static_cast<int>(std::floor(Bd.mn.y));
            if (sY < sMinY) sY = sMinY;
            if (sX < sMinX) sX = sMinX;
            if (sX >= eX || sY >= eY) continue;
            for (int iY = sY; iY < eY; ++iY)
                for (int iX = sX; iX < eX; ++iX) {
                    int index = indexT0;
                    do {
                        if (index >= numSamples) break;
                        Sample& sd = sample_at(iX, iY, index);
                        const V2 vecP = sd.pos;
                        const float time = sd.time;
                        index++;
                        st.spl_count++;
                        if (time < time0 || time > time1) continue;
                        if (vecP.x < Bd.mn.x || vecP.x > Bd.mx.x || vecP.y < Bd.mn.y || vecP.y > Bd.mx.y) continue;
                        if (F.cullableHits && Bd.mn.z > sd.occlZ) continue;
                        if (0.0f > sd.lod || sd.lod >= 1.0f) continue;
                        st.spl_bound_hits++;
                        float D;
                        V2 uv;
                        bool linear = false;
                        if (sample_motion(mp, vecP, time, D, uv, linear)) {
                            st.inside++;
                            if (!linear) st.inside_nonlinear++;
                            store_sample(mp, iX, iY, index - 1, D, uv, isOpaque);
                        }
                    } while (index < indexT1);
                }
        }
    }

    // StoreSample (bucketprocessor.cpp:1471-1569).
    void store_sample(MP* mp, int px, int py, int index, float D, V2 uv, bool isOpaque) {
        Sample& sd = sample_at(px, py, index);
        const bool isCullable = F.cullableHits;
        if (isCullable && sd.occlZ <= D) return;
        st.spl_hits++;
        pixelHasValid[(size_t)(py - F.extY0) * F.extW + (px - F.extX0)] = 1;
        const uint32_t matteFlag = mp->g->matteFlag;
        Hit* hit;
        if ((isOpaque || (matteFlag & Flag_MatteAlpha)) && isCullable) {
            hit = &sd.occl;
            if (F.midpointStore) {
                // the occluding depth becomes the second nearest opaque depth
                float hitPrevZ = FLT_MAX;
                if (hit->flags & Flag_Valid) hitPrevZ = hit->d[6];
                if (hitPrevZ < D) {
                    sd.occlZ = D;
                    if (cull) tree.set_depth(D, sd.occlIndex);
                    return;
                }
                sd.occlZ = hitPrevZ;
                if (cull) tree.set_depth(hitPrevZ, sd.occlIndex);
            } else {
                sd.occlZ = D;
                if (cull) tree.set_depth(D, sd.occlIndex);
            }
            hit->flags = Flag_Valid;
        } else {
            sd.hits.push_back(Hit());
            hit = &sd.hits.back();
            hit->flags = 0;
        }
        // InterpolateOutputs (micropolygon.cpp:1443-1462): smooth, the
        // default, bilinear over the four (here equal) vertex values;
        // constant, the first vertex's
        const Grid& g = *mp->g;
        if (g.constantShading) {
            for (int k = 0; k < 3; ++k) {
                hit->d[k] = g.Ci[k];
                hit->d[3 + k] = g.Oi[k];
            }
        } else {
            float w0 = (1 - uv.x) * (1 - uv.y);
            float w1 = uv.x * (1 - uv.y);
            float w2 = (1 - uv.x) * uv.y;
            float w3 = uv.x * uv.y;
            for (int k = 0; k < 3; ++k) {
                hit->d[k] = ((w0 * g.Ci[k] + w1 * g.Ci[k]) + w2 * g.Ci[k]) + w3 * g.Ci[k];
                hit->d[3 + k] = ((w0 * g.Oi[k] + w1 * g.Oi[k]) + w2 * g.Oi[k]) + w3 * g.Oi[k];
            }
        }
        hit->d[6] = D;
        hit->id = mp->id;
        hit->flags |= matteFlag;
    }

    // --- Bucket processing -------------------------------------------------
    // preProcess (bucketprocessor.cpp:95-206)
    void pre_process(Bucket& b) {
        const int sx = F.shiftX, sy = F.shiftY;
        int sminx = b.xPos - sx, sminy = b.yPos - sy;
        int smaxx = b.xPos + b.xSize + sx, smaxy = b.yPos + b.ySize + sy;
        if (sminx < F.cropXMin - sx) sminx = F.cropXMin - sx;
        if (sminy < F.cropYMin - sy) sminy = F.cropYMin - sy;
        if (smaxx > F.cropXMax + sx) smaxx = F.cropXMax + sx;
        if (smaxy > F.cropYMax + sy) smaxy = F.cropYMax + sy;
        // Cache segments arrive from the left and upper neighbours, which
        // were processed earlier in row-major order.
        if (b.col > F.bxMin) sminx += sx * 2;
        if (b.row > F.byMin) sminy += sy * 2;
        sMinX = sminx; sMinY = sminy; sMaxX = smaxx; sMaxY = smaxy;
        // setSamples for every sample-region pixel (imagepixel.cpp:334-359)
        const int n = F.nsamp;
        float opentime = F.o.shutter_open, closetime = F.o.shutter_close;
        for (int y = sMinY; y < sMaxY; ++y)
            for (int x = sMinX; x < sMaxX; ++x) {
                const int* shuffled;
                const V2* positions;
                const V2* dofOffsets;
                const float* times;
                const float* lods;
                if (samp.cache == 1) {
                    shuffled = &samp.shuffled[0]; positions = &samp.pos2d[0]; dofOffsets = &samp.pos2d[0];
                    times = &samp.pos1d[0]; lods = &samp.pos1d[0];
                } else {
                    shuffled = &samp.shuffled[(size_t)n * rng.rint(250)];
                    positions = &samp.pos2d[(size_t)n * rng.rint(250)];
                    dofOffsets = &samp.pos2d[(size_t)n * rng.rint(250)];
                    times = &samp.pos1d[(size_t)n * rng.rint(250)];
                    lods = &samp.pos1d[(size_t)n * rng.rint(250)];
                }
                V2 off = v2((float)x, (float)y);
                int* dofIdx = &dofIndex[((size_t)(y - F.extY0) * F.extW + (x - F.extX0)) * n];
                for (int i = 0; i < n; ++i) dofIdx[i] = shuffled[i];
                for (int i = 0; i < n; ++i) {
                    Sample& s = sample_at(x, y, i);
                    s.pos = off + positions[i];
                    s.time = (closetime - opentime) * times[i] + opentime;
                    s.lod = lods[i];
                    s.occlZ = FLT_MAX;
                    s.occl.flags = 0;
                    s.hits.clear();
                }
                // sample shuffled[i] takes lens offset i (imagepixel.cpp:357)
                for (int i = 0; i < n; ++i) {
                    const V2 d = dofOffsets[i];
                    sample_at(x, y, shuffled[i]).dofOffset =
                        project_to_circle(v2(-1.0f + 2.0f * d.x, -1.0f + 2.0f * d.y));
                }
            }
        if (cull) setup_tree();
    }

    // CqOcclusionTree::setupTree (occlusion.cpp:54-107)
    void setup_tree() {
        init_tree(tree, sMinX, sMinY, sMaxX - sMinX, sMaxY - sMinY, F.nx, F.ny);
        for (int y = sMinY; y < sMaxY; ++y)
            for (int x = sMinX; x < sMaxX; ++x)
                for (int i = 0; i < F.nsamp; ++i) {
                    int spx = F.nx * x + i % F.nx, spy = F.ny * y + i / F.nx;
                    sample_at(x, y, i).occlIndex = OcclTree::index_for_point(tree.numLevels, tree.splitXFirst,
                                                                             spx - F.nx * sMinX, spy - F.ny * sMinY);
                }
    }

    void render_waiting(Bucket& b) {
        // RenderMicroPoly (bucketprocessor.cpp:1067-1091)
        for (size_t i = 0; i < b.mps.size(); ++i) {
            if (b.mps[i]->motion) render_mp_motion(b.mps[i]);
            else if (F.dof.on) render_mp_dof(b.mps[i]);
            else render_mp_static(b.mps[i]);
        }
        for (MP* mp : b.mps) release_mp(mp);
        b.mps.clear();
        if (cull) tree.update();
    }

    // RenderSurface (bucketprocessor.cpp:942-1059); the cull test applies to
    // surfaces whose Attribute "cull" "hidden" is 1, the default (:950).
    void render_surface(Bucket& b, Surf* s) {
        if (cull && F.cullableHits && s->cachedBound && !prims[s->prim].no_cull_hidden &&
            tree.can_cull(s->rasterBound)) {
            st.occlusion_culled++;
            repost_surface(b, s);
            return;
        }
        const rh_prim& p = prims[s->prim];
        bool diceable = surf_diceable(*s, F.diceCoords, p);
        if (diceable) {
            dice_shade_bust(s);
            delete s;
        } else if (!s->fDiscard) {
            std::vector<Surf*> kids;
            surf_split(*s, kids);
            delete s;
            for (Surf* k : kids) post_surface(k);
        } else {
            delete s;
        }
    }

    // CqImagePixel::Combine without CSG (imagepixel.cpp:144-332)
    void combine_sample(Sample& sd) {
        const int depthfilter = F.o.depth_filter;
        const float* zt = F.o.zthreshold;
        if (sd.hits.empty()) {
            if (sd.occl.flags & Flag_Valid) {
                if (sd.occl.flags & Flag_Matte) {
                    // opaque matte: transparent black
                    for (int k = 0; k < 6; ++k) sd.occl.d[k] = 0.0f;
                }
                if (depthfilter == RH_DEPTH_MIDPOINT)
                    sd.occl.d[6] = (float)(0.5 * (double)(sd.occl.d[6] + sd.occlZ));
            }
            return;
        }
        if (sd.occl.flags & Flag_Valid) sd.hits.push_back(sd.occl);
        std::sort(sd.hits.begin(), sd.hits.end(),
                  [](const Hit& a, const Hit& b) { return a.d[6] < b.d[6]; });
        float col[3] = {0, 0, 0}, opa[3] = {0, 0, 0};
        float opaqueDepths[2] = {sd.occlZ, FLT_MAX};
        float maxOpaqueDepth = FLT_MAX;
        for (auto it = sd.hits.rbegin(); it != sd.hits.rend(); ++it) {
            const float* d = it->d;
            if (it->flags & Flag_Matte) {
                // lerp(t, x0, 0) = (1 - t) * x0 + t * 0 (math.h:88-91); the
                // opacity takes the hit's colour as its weight
                for (int k = 0; k < 3; ++k) {
                    const float c = col[k], o = opa[k];
                    col[k] = (1 - d[3 + k]) * c + d[3 + k] * 0.0f;
                    opa[k] = (1 - d[k]) * o + d[k] * 0.0f;
                }
            } else {
                for (int k = 0; k < 3; ++k) {
                    col[k] = col[k] * (1.0f - tclamp(d[3 + k], 0.0f, 1.0f)) + d[k];
                    opa[k] = (1.0f - opa[k]) * d[3 + k] + opa[k];
                }
            }
            if (d[3] >= zt[0] && d[4] >= zt[1] && d[5] >= zt[2]) {
                opaqueDepths[1] = opaqueDepths[0];
                opaqueDepths[0] = d[6];
                if (!(maxOpaqueDepth < FLT_MAX)) maxOpaqueDepth = d[6];
            }
        }
        Hit front = sd.hits.front();
        for (int k = 0; k < 3; ++k) { front.d[k] = col[k]; front.d[3 + k] = opa[k]; }
        front.flags |= Flag_Valid;
        float& occlDepth = front.d[6];
        if (depthfilter == RH_DEPTH_MIDPOINT) {
            occlDepth = sd.hits.size() > 1 ? (opaqueDepths[0] + opaqueDepths[1]) * 0.5f : FLT_MAX;
        } else if (depthfilter == RH_DEPTH_MAX) {
            occlDepth = maxOpaqueDepth;
        } else if (depthfilter == RH_DEPTH_AVERAGE) {
            float totDepth = 0.0f;
            int totCount = 0;
            for (const Hit& h : sd.hits)
                if (h.d[3] >= zt[0] || h.d[4] >= zt[1] || h.d[5] >= zt[2]) {
                    totDepth += h.d[6];
                    totCount++;
                }
            totDepth /= totCount;
            occlDepth = totDepth;
        } else {
            occlDepth = opaqueDepths[0];
        }
        sd.occl = front;
    }

    void post_process(Bucket& b) {
        for (int y = sMinY; y < sMaxY; ++y)
            for (int x = sMinX; x < sMaxX; ++x)
                for (int i = 0; i < F.nsamp; ++i) combine_sample(sample_at(x, y, i));
        filter_bucket(b);
        expose_bucket(b);
    }

    // FilterBucket, non-separable branch (bucketprocessor.cpp:381-744)
    void filter_bucket(Bucket& b) {
        const int xmax = F.shiftX, ymax = F.shiftY;
        const int nsub = F.nsamp;
        const int W = F.cropXMax - F.cropXMin;
        for (int y = b.yPos; y < b.yPos + b.ySize; ++y)
            for (int x = b.xPos; x < b.xPos + b.xSize; ++x) {
                if (x < F.cropXMin || x >= F.cropXMax || y < F.cropYMin || y >= F.cropYMax) continue;
                float ycent = y + 0.5f, xcent = x + 0.5f;
                float gTot = 0.0f;
                int SampleCount = 0;
                float samplesv[7] = {0, 0, 0, 0, 0, 0, 0};
                for (int fy = -ymax; fy <= ymax; ++fy)
                    for (int fx = -xmax; fx <= xmax; ++fx) {
                        int index = ((fy + ymax) * (2 * xmax + 1) + fx + xmax) * nsub;
                        int si = 0;
                        for (int sy = 0; sy < F.ny; ++sy)
                            for (int sx = 0; sx < F.nx; ++sx, ++si) {
                                Sample& sd = sample_at(x + fx, y + fy, si);
                                V2 vecS = sd.pos - v2(xcent, ycent);
                                if (vecS.x >= -F.xfwo2 && vecS.y >= -F.yfwo2 && vecS.x <= F.xfwo2 && vecS.y <= F.yfwo2) {
                                    float g = F.filterValues[index + si];
                                    gTot += g;
                                    if (sd.occl.flags & Flag_Valid) {
                                        for (int k = 0; k < 7; ++k) samplesv[k] += sd.occl.d[k] * g;
                                        SampleCount++;
                                    }
                                }
                            }
                    }
                float* out = &channels[((size_t)(y - F.cropYMin) * W + (x - F.cropXMin)) * 9];
                float coverage;
                if (SampleCount == 0) {
                    for (int k = 0; k < 9; ++k) out[k] = 0.0f;
                    out[7] = FLT_MAX;
                    coverage = 0.0f;
                } else {
                    float oneOverGTot = (float)(1.0 / gTot);
                    for (int k = 0; k < 6; ++k) out[k] = samplesv[k] * oneOverGTot;
                    out[7] = samplesv[6] * oneOverGTot;
                    coverage = (SampleCount >= nsub) ? 1.0f : (float)SampleCount / (float)nsub;
                }
                float a = (out[3] + out[4] + out[5]) / 3.0f;
                out[6] = a * coverage;
                out[8] = coverage;
            }
    }

    // ExposeBucket (bucketprocessor.cpp:766-806)
    void expose_bucket(Bucket& b) {
        float gain = F.o.exposure_gain, gamma = F.o.exposure_gamma;
        if (gain == 1.0 && gamma == 1.0) return;
        float oneovergamma = 1.0f / gamma;
        const int W = F.cropXMax - F.cropXMin;
        for (int y = b.yPos; y < b.yPos + b.ySize; ++y)
            for (int x = b.xPos; x < b.xPos + b.xSize; ++x) {
                if (x < F.cropXMin || x >= F.cropXMax || y < F.cropYMin || y >= F.cropYMax) continue;
                float* out = &channels[((size_t)(y - F.cropYMin) * W + (x - F.cropXMin)) * 9];
                for (int k = 0; k < 3; ++k) {
                    if (gain != 1.0) out[k] *= gain;
                    if (gamma != 1.0) out[k] = (float)std::pow((double)out[k], (double)oneovergamma);
                }
            }
    }

    // CqDDManager::DisplayBucket + FormatBucketForDisplay (ddmanager.cpp:146-171,
    // 1022-1118): one RandomFloat per pixel per display, rgba 8-bit.
    void display_bucket(Bucket& b) {
        int xmin = b.xPos, ymin = b.yPos, xmaxp1 = b.xPos + b.xSize, ymaxp1 = b.yPos + b.ySize;
        if (xmaxp1 <= F.cropXMin || ymaxp1 <= F.cropYMin || xmin > F.cropXMax || ymin > F.cropYMax) return;
        const int W = F.cropXMax - F.cropXMin;
        const rh_options& o = F.o;
        const float zero = 0.0f;
        for (int d = 0; d < o.num_displays; ++d)
            for (int y = b.yPos; y < b.yPos + b.ySize; ++y)
                for (int x = b.xPos; x < b.xPos + b.xSize; ++x) {
                    double s = rng.rfloat();
                    if (x < F.cropXMin || x >= F.cropXMax || y < F.cropYMin || y >= F.cropYMax) continue;
                    if (d == ditherIndex) ditherRec[(size_t)(y - F.cropYMin) * W + (x - F.cropXMin)] = (float)s;
                    if (d != 0) continue;
                    const float* ch = &channels[((size_t)(y - F.cropYMin) * W + (x - F.cropXMin)) * 9];
                    const float vals[4] = {ch[0], ch[1], ch[2], ch[6]};
                    uint8_t* out = &rgba8[((size_t)(y - F.cropYMin) * W + (x - F.cropXMin)) * 4];
                    for (int k = 0; k < 4; ++k) {
                        double value = vals[k];
                        if (o.quantize_one != 0) {
                            value = (double)lround(zero + value * (double)(o.quantize_one - zero) +
                                                   ((double)o.quantize_dither * s));
                            value = tclamp<double>(value, o.quantize_min, o.quantize_max);
                        }
                        out[k] = static_cast<uint8_t>(value);
                    }
                }
    }

    void render() {
        const rh_options& o = F.o;
        // WorldBegin reseeds the global stream (ri.cpp:660); the sampler is
        // built in RenderImage (imagebuffer.cpp:694-704).
        rng.reseed('a' + 'q' + 's' + 'i' + 's');
        if (o.jitter) build_multijitter(rng, F.nx, F.ny, samp);
        else build_grid_sampler(F.nx, F.ny, samp);
        for (int r = F.byMin; r < F.byMax; ++r)
            for (int c = F.bxMin; c < F.bxMax; ++c) {
                Bucket& b = bucket(c, r);
                pre_process(b);
                render_waiting(b);
                while (!b.heap.empty()) {
                    Surf* s = b.heap.front();
                    std::pop_heap(b.heap.begin(), b.heap.end(), ClosestSurface());
                    b.heap.pop_back();
                    render_surface(b, s);
                    render_waiting(b);
                }
                render_waiting(b);
                post_process(b);
                for (Surf* s : b.heap) delete s;
                b.heap.clear();
                b.processed = true;
                display_bucket(b);
            }
    }
};

}  // namespace orc

using namespace orc;

extern "C" {

int oracle_frame_size(const rh_options* opt, int32_t* w, int32_t* h, int32_t* ew, int32_t* eh, int32_t* spp) {
    Frame F;
    setup_frame(*opt, F);
    if (w) *w = F.cropXMax - F.cropXMin;
    if (h) *h = F.cropYMax - F.cropYMin;
    if (ew) *ew = F.extW;
    if (eh) *eh = F.extH;
    if (spp) *spp = F.nsamp;
    return 0;
}

int oracle_render(const rh_options* opt, const rh_prim* prims, int32_t nprims, int32_t occlusion_cull,
                  float* channels, uint8_t* rgba8, uint32_t* mp_id, float* depth, float* rgbaz,
                  rh_stats* stats) {
    std::unique_ptr<Hider> H(new Hider());
    memset(&H->st, 0, sizeof(H->st));
    setup_frame(*opt, H->F);
    Frame& F = H->F;
    H->prims = prims;
    H->cull = occlusion_cull != 0;
    // Depth of field together with motion blur (the motion Sample's DoF
    // branch, micropolygon.cpp:1811-1870) is not restated.
    if (F.dof.on)
        for (int32_t i = 0; i < nprims; ++i)
            if (prims[i].nkeys != 1) return -2;
    g_dof = &F.dof;
    struct DofReset { ~DofReset() { g_dof = nullptr; } } dofReset;
    H->samples.resize((size_t)F.extW * F.extH * F.nsamp);
    H->dofIndex.assign((size_t)F.extW * F.extH * F.nsamp, 0);
    for (auto& s : H->samples) {
        s.occlZ = FLT_MAX; s.occl.flags = 0; s.occlIndex = 0; s.pos = v2(0, 0); s.dofOffset = v2(0, 0);
    }
    H->pixelHasValid.assign((size_t)F.extW * F.extH, 0);
    const int W = F.cropXMax - F.cropXMin, Hh = F.cropYMax - F.cropYMin;
    H->channels.assign((size_t)W * Hh * 9, 0.0f);
    H->rgba8.assign((size_t)W * Hh * 4, 0);
    H->ditherRec.assign((size_t)W * Hh, 0.0f);
    H->set_image();
    // Timed frame: posting (PostSurface of every primitive) + RenderImage.
    auto t0 = std::chrono::steady_clock::now();
    // Posting order = primitive order (RiSphere/RiPatch -> StorePrimitive -> PostSurface)
    for (int i = 0; i < nprims; ++i) {
        const rh_prim& p = prims[i];
        Surf* s = new Surf();
        s->type = p.type;
        s->prim = i;
        if (p.type == RH_PRIM_SPHERE) {
            // CqSphere constructor (quadrics.cpp:352-366), thetamin = 0
            s->radius = p.sphere[0];
            s->thetaMin = 0.0f;
            s->thetaMax = p.sphere[3];
            s->phiMin = (float)std::asin((double)(p.sphere[1] / s->radius));
            s->phiMax = (float)std::asin((double)(p.sphere[2] / s->radius));
            for (int key = 0; key < p.nkeys; ++key) {
                Mat M;
                bool ident = true;
                for (int r = 0; r < 4; ++r)
                    for (int c = 0; c < 4; ++c) {
                        M.m[r][c] = p.xform[key][4 * r + c];
                        if (M.m[r][c] != (r == c ? 1.0f : 0.0f)) ident = false;
                    }
                M.ident = ident;
                (key == 0 ? s->matTx : s->matTx1) = M;
            }
        } else {
            int n = p.type == RH_PRIM_BICUBIC ? 16 : 4;
            for (int k = 0; k < n; ++k) s->cp[k] = v3(p.P[k][0], p.P[k][1], p.P[k][2]);
            for (int k = 0; k < n; ++k) s->cp1[k] = v3(p.P1[k][0], p.P1[k][1], p.P1[k][2]);
        }
        if (p.nkeys == 2) {
            s->motion = true;
            s->keyTime[0] = p.key_time[0];
            s->keyTime[1] = p.key_time[1];
        }
        H->post_surface(s);
    }
    auto t1 = std::chrono::steady_clock::now();
    H->render();
    auto t2 = std::chrono::steady_clock::now();
    H->st.ms_schedule = std::chrono::duration<double, std::milli>(t1 - t0).count();
    H->st.ms_device = std::chrono::duration<double, std::milli>(t2 - t0).count();
    if (channels) memcpy(channels, H->channels.data(), H->channels.size() * sizeof(float));
    if (rgba8) memcpy(rgba8, H->rgba8.data(), H->rgba8.size());
    const size_t ns = H->samples.size();
    for (size_t i = 0; i < ns; ++i) {
        const Sample& s = H->samples[i];
        bool valid = (s.occl.flags & Flag_Valid) != 0;
        if (mp_id) mp_id[i] = valid ? s.occl.id : 0xffffffffu;
        if (depth) {
            // depth of the front-most hit: the occluding hit's depth after
            // the combine for opaque-only samples (its own, or under the
            // midpoint filter the mean with the occluding depth), the first
            // sorted hit otherwise.
            float z = FLT_MAX;
            if (valid) z = s.hits.empty() ? s.occl.d[6] : s.hits.front().d[6];
            depth[i] = z;
        }
        if (rgbaz) {
            for (int k = 0; k < 7; ++k) rgbaz[i * 7 + k] = valid ? s.occl.d[k] : (k == 6 ? FLT_MAX : 0.0f);
        }
    }
    if (stats) *stats = H->st;
    return 0;
}

int oracle_dice(const rh_prim* p, int32_t u, int32_t v, float* out) {
    Surf s;
    s.type = p->type;
    if (p->type == RH_PRIM_SPHERE) {
        s.radius = p->sphere[0];
        s.thetaMin = 0.0f;
        s.thetaMax = p->sphere[3];
        s.phiMin = (float)std::asin((double)(p->sphere[1] / s.radius));
        s.phiMax = (float)std::asin((double)(p->sphere[2] / s.radius));
        bool ident = true;
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) {
                s.matTx.m[r][c] = p->xform[0][4 * r + c];
                if (s.matTx.m[r][c] != (r == c ? 1.0f : 0.0f)) ident = false;
            }
        s.matTx.ident = ident;
    } else {
        int n = p->type == RH_PRIM_BICUBIC ? 16 : 4;
        for (int k = 0; k < n; ++k) s.cp[k] = v3(p->P[k][0], p->P[k][1], p->P[k][2]);
    }
    s.uDice = u;
    s.vDice = v;
    Grid g;
    surf_dice(s, g);
    for (size_t i = 0; i < g.P.size(); ++i) { out[3 * i] = g.P[i].x; out[3 * i + 1] = g.P[i].y; out[3 * i + 2] = g.P[i].z; }
    return 0;
}

int oracle_sample_positions_of(const rh_options* opt, int32_t index, float* xy, float* dither);
int oracle_sample_positions(const rh_options* opt, float* xy, float* dither) {
    return oracle_sample_positions_of(opt, 0, xy, dither);
}

// The same with the dither draws of display request `index` (one
// RandomFloat per display pixel per request, the requests in order within
// each bucket: ddmanager.cpp:146-171, 1050).
int oracle_sample_positions_of(const rh_options* opt, int32_t index, float* xy, float* dither) {
    std::unique_ptr<Hider> H(new Hider());
    H->ditherIndex = index;
    memset(&H->st, 0, sizeof(H->st));
    setup_frame(*opt, H->F);
    Frame& F = H->F;
    H->prims = nullptr;
    H->cull = false;
    H->samples.resize((size_t)F.extW * F.extH * F.nsamp);
    H->dofIndex.assign((size_t)F.extW * F.extH * F.nsamp, 0);
    H->pixelHasValid.assign((size_t)F.extW * F.extH, 0);
    const int W = F.cropXMax - F.cropXMin, Hh = F.cropYMax - F.cropYMin;
    H->channels.assign((size_t)W * Hh * 9, 0.0f);
    H->rgba8.assign((size_t)W * Hh * 4, 0);
    H->ditherRec.assign((size_t)W * Hh, 0.0f);
    H->set_image();
    H->render();
    if (xy)
        for (size_t i = 0; i < H->samples.size(); ++i) { xy[2 * i] = H->samples[i].pos.x; xy[2 * i + 1] = H->samples[i].pos.y; }
    if (dither) memcpy(dither, H->ditherRec.data(), H->ditherRec.size() * sizeof(float));
    return 0;
}

// Per-sample shutter times of the extended image (setSamples,
// imagepixel.cpp:334-359), for tests.
int oracle_sample_times(const rh_options* opt, float* times) {
    std::unique_ptr<Hider> H(new Hider());
    memset(&H->st, 0, sizeof(H->st));
    setup_frame(*opt, H->F);
    Frame& F = H->F;
    H->prims = nullptr;
    H->cull = false;
    H->samples.resize((size_t)F.extW * F.extH * F.nsamp);
    H->dofIndex.assign((size_t)F.extW * F.extH * F.nsamp, 0);
    H->pixelHasValid.assign((size_t)F.extW * F.extH, 0);
    const int W = F.cropXMax - F.cropXMin, Hh = F.cropYMax - F.cropYMin;
    H->channels.assign((size_t)W * Hh * 9, 0.0f);
    H->rgba8.assign((size_t)W * Hh * 4, 0);
    H->ditherRec.assign((size_t)W * Hh, 0.0f);
    H->set_image();
    H->render();
    for (size_t i = 0; i < H->samples.size(); ++i) times[i] = H->samples[i].time;
    return 0;
}

// Per-sample lens offsets of the extended image (setSamples,
// imagepixel.cpp:338-357): 2 floats per sample, in the unit disk, and the
// lens cell of each sample.
int oracle_sample_lens(const rh_options* opt, float* xy, int32_t* cell) {
    std::unique_ptr<Hider> H(new Hider());
    memset(&H->st, 0, sizeof(H->st));
    setup_frame(*opt, H->F);
    Frame& F = H->F;
    H->prims = nullptr;
    H->cull = false;
    H->samples.resize((size_t)F.extW * F.extH * F.nsamp);
    H->dofIndex.assign((size_t)F.extW * F.extH * F.nsamp, 0);
    H->pixelHasValid.assign((size_t)F.extW * F.extH, 0);
    const int W = F.cropXMax - F.cropXMin, Hh = F.cropYMax - F.cropYMin;
    H->channels.assign((size_t)W * Hh * 9, 0.0f);
    H->rgba8.assign((size_t)W * Hh * 4, 0);
    H->ditherRec.assign((size_t)W * Hh, 0.0f);
    H->set_image();
    H->render();
    for (size_t i = 0; i < H->samples.size(); ++i) {
        xy[2 * i] = H->samples[i].dofOffset.x;
        xy[2 * i + 1] = H->samples[i].dofOffset.y;
    }
    // the lens cell through which each sample is visited (GetDofOffsetIndex inverted)
    if (cell)
        for (size_t px = 0; px < (size_t)F.extW * F.extH; ++px)
            for (int k = 0; k < F.nsamp; ++k) cell[px * F.nsamp + H->dofIndex[px * F.nsamp + k]] = k;
    return 0;
}

// The frame's circle of confusion (x and y raster radii) at n camera
// depths, its lens-cell bounds (4 floats per cell: min x, min y, max x,
// max y) and MinCoCForBound of [z[i], z[i+1]] for i < n - 1.
int oracle_dof(const rh_options* opt, int32_t n, const float* z, float* coc, float* cells, float* minCoc) {
    Frame F;
    setup_frame(*opt, F);
    if (!F.dof.on) return -1;
    for (int i = 0; i < n; ++i) {
        V2 c = circle_of_confusion(F.dof, z[i]);
        coc[2 * i] = c.x;
        coc[2 * i + 1] = c.y;
        if (i + 1 < n) minCoc[i] = min_coc_for_bound(F.dof, z[i], z[i + 1]);
    }
    for (int k = 0; k < F.nsamp; ++k) {
        cells[4 * k] = F.dofBounds[k].mn.x; cells[4 * k + 1] = F.dofBounds[k].mn.y;
        cells[4 * k + 2] = F.dofBounds[k].mx.x; cells[4 * k + 3] = F.dofBounds[k].mx.y;
    }
    return 0;
}

void oracle_mt_outputs(uint32_t seed, int32_t n, uint32_t* out) {
    Rng r;
    r.reseed(seed);
    for (int i = 0; i < n; ++i) out[i] = r.next();
}

void oracle_random_floats(uint32_t seed, int32_t n, float* out) {
    Rng r;
    r.reseed(seed);
    for (int i = 0; i < n; ++i) out[i] = r.rfloat();
}

void oracle_round_f(const float* x, int32_t n, int64_t* fl, int64_t* ce, int64_t* ro) {
    for (int i = 0; i < n; ++i) {
        if (fl) fl[i] = lfloor(x[i]);
        if (ce) ce[i] = lceil(x[i]);
        if (ro) ro[i] = lround(x[i]);
    }
}

void oracle_inv_bilinear(const float* abcd, const float* p, int32_t n, float* uv) {
    Hider::HitCache c;
    V3 pin[4];
    for (int i = 0; i < 4; ++i) pin[i] = v3(abcd[2 * i], abcd[2 * i + 1], 0);
    Hider::cache_point_in_poly(c, pin, 0 | (1 << 2) | (3 << 4) | (2 << 6));
    for (int i = 0; i < n; ++i) {
        V2 r = Hider::inv_bilinear(c, v2(p[2 * i], p[2 * i + 1]));
        uv[2 * i] = r.x;
        uv[2 * i + 1] = r.y;
    }
}

void oracle_bilerp2(const float* abcd, const float* uv, int32_t n, float* out) {
    for (int i = 0; i < n; ++i) {
        V2 t = v2(uv[2 * i], uv[2 * i + 1]);
        out[2 * i] = Hider::bilerp(abcd[0], abcd[2], abcd[4], abcd[6], t);
        out[2 * i + 1] = Hider::bilerp(abcd[1], abcd[3], abcd[5], abcd[7], t);
    }
}

int32_t oracle_tree_index(int32_t numLevels, int32_t splitXFirst, int32_t x, int32_t y) {
    return OcclTree::index_for_point(numLevels, splitXFirst != 0, x, y);
}

int32_t oracle_can_cull(int32_t sMinX, int32_t sMinY, int32_t w, int32_t h, int32_t xS, int32_t yS,
                        const float* cellDepth, const float* bound) {
    OcclTree t;
    init_tree(t, sMinX, sMinY, w, h, xS, yS);
    for (int y = 0; y < h * yS; ++y)
        for (int x = 0; x < w * xS; ++x)
            t.depth[OcclTree::index_for_point(t.numLevels, t.splitXFirst, x, y)] = cellDepth[(size_t)y * w * xS + x];
    t.propagate();
    Bound b;
    b.mn = v3(bound[0], bound[1], bound[4]);
    b.mx = v3(bound[2], bound[3], bound[5]);
    return t.can_cull(b) ? 1 : 0;
}

float oracle_filter(int32_t filter, float x, float y, float xw, float yw) {
    return filter_eval(filter, x, y, xw, yw);
}

void oracle_jitter_tables(int32_t nx, int32_t ny, float* pos2d, float* pos1d, int32_t* shuffled) {
    Rng r;
    r.reseed('a' + 'q' + 's' + 'i' + 's');
    SamplerTables t;
    build_multijitter(r, nx, ny, t);
    for (size_t i = 0; i < t.pos2d.size(); ++i) {
        if (pos2d) { pos2d[2 * i] = t.pos2d[i].x; pos2d[2 * i + 1] = t.pos2d[i].y; }
        if (pos1d) pos1d[i] = t.pos1d[i];
        if (shuffled) shuffled[i] = t.shuffled[i];
    }
}

int32_t oracle_mp_sample(const float* P, float x, float y, float* depth) {
    Grid g;
    g.cu = 1; g.cv = 1;
    g.P.resize(4);
    for (int i = 0; i < 4; ++i) g.P[i] = v3(P[3 * i], P[3 * i + 1], P[3 * i + 2]);
    int code = Hider::vertex_order(g, 0);
    Hider::HitCache c;
    V3 pin[4] = {g.P[0], g.P[1], g.P[2], g.P[3]};
    Hider::cache_point_in_poly(c, pin, code);
    float D = 0;
    V2 uv;
    if (Hider::contains(c, v2(x, y), D, uv)) { if (depth) *depth = D; return 1; }
    return 0;
}

// selectDataFormat (ddmanager.cpp:249-283) for the rgba request: the
// PkDspy* code of the values a display is given (1 float32, 2 uint32,
// 3 int32, 4 uint16, 5 int16, 6 uint8, 7 int8).
int32_t oracle_display_type(const rh_options* o) {
    if (o->quantize_one == 0) return 1;
    if (o->quantize_min >= 0) return o->quantize_max <= 255 ? 6 : (o->quantize_max <= 65535 ? 4 : 2);
    if (o->quantize_min >= -128 && o->quantize_max <= 127) return 7;
    if (o->quantize_min >= -32768 && o->quantize_max <= 32767) return 5;
    return 3;
}

// FormatBucketForDisplay's per-value path (ddmanager.cpp:1045-1118) for one
// value: quantised with (one, min, max, dither) q and draw s when one != 0,
// stored as `type` with the reference's own static_casts (x86-64 g++ code,
// as its build); returns the next byte.
static unsigned char* oracle_put_value(const float* q, double value, double s, int32_t type, unsigned char* pdata) {
            const float zero = 0.0f;
            if (q[0] != 0) {
                // Aqsis::lround (floor(x - 0.5) + 1), not the C library's
                value = (double)orc::lround(zero + value * (double)(q[0] - zero) + ((double)q[3] * s));
                value = tclamp<double>(value, q[1], q[2]);
            }
            switch (type) {
            case 1: { float v = (float)value; memcpy(pdata, &v, 4); pdata += 4; break; }
            case 2: {
                value = tclamp<double>(value, 0, (double)std::numeric_limits<uint32_t>::max());
                uint32_t v = static_cast<uint32_t>(value);
                memcpy(pdata, &v, 4); pdata += 4; break;
            }
            case 3: {
                value = tclamp<double>(value, (double)std::numeric_limits<int32_t>::min(),
                                       (double)std::numeric_limits<int32_t>::max());
                int32_t v = static_cast<int32_t>(value);
                memcpy(pdata, &v, 4); pdata += 4; break;
            }
            case 4: { uint16_t v = static_cast<uint16_t>(value); memcpy(pdata, &v, 2); pdata += 2; break; }
            case 5: { int16_t v = static_cast<int16_t>(value); memcpy(pdata, &v, 2); pdata += 2; break; }
            case 6: { uint8_t v = static_cast<uint8_t>(value); *pdata++ = v; break; }
            default: { int8_t v = static_cast<int8_t>(value); memcpy(pdata, &v, 1); pdata += 1; break; }
            }
            return pdata;
}

// FormatBucketForDisplay over npix pixels: r, g, b = channels 0..2 and a =
// channel 6 of each pixel's 9, with the options' Quantize "rgba" settings.
void oracle_display_values(const rh_options* o, const float* channels, const float* dither, int64_t npix,
                           int32_t type, void* out) {
    unsigned char* pdata = static_cast<unsigned char*>(out);
    const float q[4] = {o->quantize_one, o->quantize_min, o->quantize_max, o->quantize_dither};
    for (int64_t i = 0; i < npix; ++i) {
        const float* ch = channels + i * 9;
        const float vals[4] = {ch[0], ch[1], ch[2], ch[6]};
        for (int k = 0; k < 4; ++k) pdata = oracle_put_value(q, vals[k], dither[i], type, pdata);
    }
}

// A display request of mode bits (1 rgb, 2 a, 4 z) with Quantize settings
// q: its channels in the order it builds its formats (ddmanager.cpp:456-483):
// a (channel 6), then r, g, b (0..2), then z (the filtered depth, channel 7).
void oracle_display_request_values(const float* q, int32_t mode, const float* channels, const float* dither,
                                   int64_t npix, int32_t type, void* out) {
    unsigned char* pdata = static_cast<unsigned char*>(out);
    for (int64_t i = 0; i < npix; ++i) {
        const float* ch = channels + i * 9;
        if (mode & 2) pdata = oracle_put_value(q, ch[6], dither[i], type, pdata);
        if (mode & 1)
            for (int k = 0; k < 3; ++k) pdata = oracle_put_value(q, ch[k], dither[i], type, pdata);
        if (mode & 4) pdata = oracle_put_value(q, ch[7], dither[i], type, pdata);
    }
}

}  // extern "C"
