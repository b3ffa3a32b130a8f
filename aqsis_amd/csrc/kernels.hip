// HIP kernels of the hider for gfx950 (CDNA4).
//
//   rh_dice_bust        one workgroup per grid: dice (sphere / bicubic /
//                       bilinear), project to raster, and write one hit-test
//                       record per micropolygon (edge equations, inverse-
//                       bilinear set-up, depths, bound, opacity).
//   rh_bin_count/scan/write
//                       per (bucket, candidate grid) segment: keep the
//                       micropolygons the bucket would have received (AddMPG)
//                       and append them, in order, to the lists of the tasks
//                       (parts of the bucket's sample region) they touch.
//   rh_plan_items       batched form: cut tasks into work items by cost,
//                       largest first.
//   rh_sample           one wave per block of 64 samples, a workgroup per
//                       work item: the reference's per-sample tests (bound,
//                       occlusion, LOD, point-in-quad, inverse bilinear,
//                       depth, store) in list order; transparent hits go to
//                       per-sample lists.
//   rh_combine          samples with transparent hits: depth sort as
//                       std::sort, back-to-front compositing.
//   rh_box_masks        per frame shape: each sample's filter-box tests.
//   rh_filter_tiled     8x8-pixel tiles, halo in LDS: the filter gather,
//                       exposure, quantise + dither (rh_filter: the general
//                       thread-per-pixel form).
//
// All single-precision code is built with -ffp-contract=off and correctly
// rounded division; every expression keeps the operand order of the aqsis
// code it cites.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cstdint>

#include "devtypes.h"
#include "hitsort.h"
#include "cullgeom.h"

namespace {

__device__ __forceinline__ int floor_i(float x) {  // static_cast<int>(std::floor(x))
    return (int)floorf(x);
}
__device__ __forceinline__ long lceil_f(float x) {  // math.h:57-61
    long t = (long)x;
    return t + (x > 0 && x != (float)t);
}
// (int)lceil(x), the low 32 bits of lceil's long: ceil(x) inside
// (-2^31, 2^31), the long form elsewhere and for NaN.
__device__ __forceinline__ int lceil_i(float x) {
    if (fabsf(x) < 2147483648.0f) return (int)ceilf(x);
    return (int)lceil_f(x);
}
__device__ __forceinline__ long lfloor_d(double x) {  // math.h:47-51
    // static_cast<long> as x86-64 performs it (cvttsd2si): NaN and values
    // outside the long range give LONG_MIN; the subtraction wraps.  So a NaN
    // colour (pow of a negative filtered value) quantises to the minimum.
    const long t = (x >= -9223372036854775808.0 && x < 9223372036854775808.0) ? (long)x : (long)(1ull << 63);
    return (long)((unsigned long)t - (unsigned long)(x < 0 && x != (double)t));
}
// lfloor / lceil of math.h:47-61 to int where the result fits.
__device__ __forceinline__ int lfloor_i(float x) {
    if (fabsf(x) < 2147483648.0f) return (int)floorf(x);
    long t = (long)x;
    return (int)(t - (x < 0 && x != (float)t));
}
__device__ __forceinline__ float fmin_ref(float a, float b) { return (a < b) ? a : b; }
__device__ __forceinline__ float fmax_ref(float a, float b) { return (a < b) ? b : a; }

struct f3d {
    float x, y, z;
};
__device__ __forceinline__ f3d mk3(float x, float y, float z) { f3d r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ f3d add3(f3d a, f3d b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3d mul3(f3d a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }

// Moving micropolygons.  Key bounds as CqMovingMicroPolygonKey::GetBound
// (micropolygon.cpp:1974-1989), kb[k] = {min x, min y, min z, max x, max y,
// max z} of key k's four vertices.
__device__ __forceinline__ void mb_key_bound(const f3d* q, float* kb) {
    kb[0] = fmin_ref(q[0].x, fmin_ref(q[1].x, fmin_ref(q[2].x, q[3].x)));
    kb[1] = fmin_ref(q[0].y, fmin_ref(q[1].y, fmin_ref(q[2].y, q[3].y)));
    kb[2] = fmin_ref(q[0].z, fmin_ref(q[1].z, fmin_ref(q[2].z, q[3].z)));
    kb[3] = fmax_ref(q[0].x, fmax_ref(q[1].x, fmax_ref(q[2].x, q[3].x)));
    kb[4] = fmax_ref(q[0].y, fmax_ref(q[1].y, fmax_ref(q[2].y, q[3].y)));
    kb[5] = fmax_ref(q[0].z, fmax_ref(q[1].z, fmax_ref(q[2].z, q[3].z)));
}

// CqMicroPolygonMotion::BuildBoundList's sub-bound count
// (micropolygon.cpp:1697-1712): the raster move of the first vertex in key
// 0's bound widths, at least 1, at most timeRanges.  lfloor of an infinite
// or NaN ratio is LONG_MIN on x86-64, so the count is then 1.
__device__ __forceinline__ int mb_divisions(const float* kb0, f3d a0, f3d a1, int timeRanges) {
    const float wx = kb0[3] - kb0[0], wy = kb0[4] - kb0[1];
    const float polyLen2 = (wy == 0.0f) ? wx * wx : ((wx == 0.0f) ? wy * wy : (wx * wx) + (wy * wy));
    const float mxd = a0.x - a1.x, myd = a0.y - a1.y;
    const float moveDist2 = (myd == 0.0f) ? mxd * mxd : ((mxd == 0.0f) ? myd * myd : (mxd * mxd) + (myd * myd));
    const float lengths = sqrtf(moveDist2 / polyLen2);
    int moved = (lengths < 9.2e18f) ? (lengths < 2147483648.0f ? lfloor_i(lengths) : 0x7fffffff) : 1;
    if (moved < 1) moved = 1;
    return moved < timeRanges ? moved : timeRanges;
}

// The key bounds mixed at `mix` (extrapolated outside [0, 1]), component c.
__device__ __forceinline__ float mb_mix(const float (*kb)[6], int c, float mix) {
    return kb[0][c] + mix * (kb[1][c] - kb[0][c]);
}

// The raster hull {min x, max x, min y, max y} of every sub-bound
// BuildBoundList (micropolygon.cpp:1689-1756) makes: key 0's bound and the
// mixed bounds at each sub-bound's end time, with the same time steps.
__device__ __forceinline__ float4 mb_hull(const float (*kb)[6], int divisions, float T0, float T1, float opentime,
                                          float closetime) {
    const float dt = (closetime - opentime) / divisions;
    float time = opentime + dt;
    float4 h = make_float4(kb[0][0], kb[0][3], kb[0][1], kb[0][4]);
    for (int i = 0; i < divisions; ++i) {
        const float mix = (time - T0) / (T1 - T0);
        h.x = fmin_ref(h.x, mb_mix(kb, 0, mix));
        h.y = fmax_ref(h.y, mb_mix(kb, 3, mix));
        h.z = fmin_ref(h.z, mb_mix(kb, 1, mix));
        h.w = fmax_ref(h.w, mb_mix(kb, 4, mix));
        time += dt;
    }
    return h;
}

// CqMatrix * CqVector3D (matrix.h:717-750); m[16] is [row][col] row-major.
__device__ __forceinline__ f3d xform_point(const float* m, bool ident, f3d v) {
    if (ident) return v;
    float h = m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15];
    f3d r;
    r.x = m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12];
    r.y = m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13];
    r.z = m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14];
    if (h != 1) {
        float inv = 1 / h;
        r.x = r.x * inv; r.y = r.y * inv; r.z = r.z * inv;
    }
    return r;
}

// CqForwardDiffBezier (forwarddiff.h:35-103)
struct FD {
    float M00, M01, M02, M03, M10, M11, M12, M13, M20, M21, M22, M23;
    f3d f, df, ddf, dddf;
    __device__ void init(float dt) {
        float dt2 = dt * dt, dt3 = dt2 * dt, dt3_3 = 3 * dt3, dt3_6 = 6 * dt3, dt3_18 = 18 * dt3, dt2_3 = 3 * dt2,
              dt2_6 = 6 * dt2, dt_3 = 3 * dt;
        M00 = -dt3_6; M01 = dt3_18; M02 = -dt3_18; M03 = dt3_6;
        M10 = dt2_6 - dt3_6; M11 = dt3_18 - 2 * dt2_6; M12 = dt2_6 - dt3_18; M13 = dt3_6;
        M20 = dt2_3 - dt_3 - dt3; M21 = dt3_3 - dt2_6 + dt_3; M22 = dt2_3 - dt3_3; M23 = dt3;
    }
    __device__ void calc(f3d A, f3d B, f3d C, f3d D) {
        f = A;
        df = add3(add3(add3(mul3(A, M20), mul3(B, M21)), mul3(C, M22)), mul3(D, M23));
        ddf = add3(add3(add3(mul3(A, M10), mul3(B, M11)), mul3(C, M12)), mul3(D, M13));
        dddf = add3(add3(add3(mul3(A, M00), mul3(B, M01)), mul3(C, M02)), mul3(D, M03));
    }
    __device__ f3d get() {
        f3d r = f;
        f = add3(f, df);
        df = add3(df, ddf);
        ddf = add3(ddf, dddf);
        return r;
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// Dice + bust
// ---------------------------------------------------------------------------
// One key's grid points in camera space into V (LDS): the dicers of the
// three surface types.  `type` without the motion flag; prm: the key's
// parameters.
__device__ __forceinline__ void dice_points(int type, int cu, int cv, const float* __restrict__ prm, float4* V,
                                            float* rows) {
    const int tid = threadIdx.x;
    const int nu = cu + 1, nv = (cu + 1) * (cv + 1);
    if (type == RH_PRIM_SPHERE) {
        // CqSphere::DicePoints + DiceAll (quadrics.cpp:457-492, 82-254)
        const float radius = prm[SPH_RADIUS_D];
        const bool ident = prm[SPH_IDENT_D] != 0.0f;
        const float* sinT = prm + SPH_TABLES_D;
        const float* cosT = sinT + nu;
        const float* sinP = cosT + nu;
        const float* cosP = sinP + (cv + 1);
        for (int i = tid; i < nv; i += blockDim.x) {
            int u = i % nu, v = i / nu;
            f3d unitP = mk3(cosT[u] * cosP[v], sinT[u] * cosP[v], sinP[v]);
            f3d p = mk3(radius * unitP.x, radius * unitP.y, radius * unitP.z);
            p = xform_point(prm + SPH_MAT_D, ident, p);
            V[i] = make_float4(p.x, p.y, p.z, 0.0f);
        }
    } else if (type == RH_PRIM_BICUBIC) {
        // bicubicPatchNatDice (patch.cpp:205-243): four v-direction forward
        // differences give each row's control points, then one u-direction
        // forward difference per row.
        if (tid < 4) {
            FD fd;
            fd.init(1.0f / (float)cv);
            const float* c = prm;
            fd.calc(mk3(c[3 * tid], c[3 * tid + 1], c[3 * tid + 2]),
                    mk3(c[3 * (4 + tid)], c[3 * (4 + tid) + 1], c[3 * (4 + tid) + 2]),
                    mk3(c[3 * (8 + tid)], c[3 * (8 + tid) + 1], c[3 * (8 + tid) + 2]),
                    mk3(c[3 * (12 + tid)], c[3 * (12 + tid) + 1], c[3 * (12 + tid) + 2]));
            for (int iv = 0; iv <= cv; ++iv) {
                f3d q = fd.get();
                float* r = rows + (iv * 4 + tid) * 3;
                r[0] = q.x; r[1] = q.y; r[2] = q.z;
            }
        }
        __syncthreads();
        for (int iv = tid; iv <= cv; iv += blockDim.x) {
            FD fd;
            fd.init(1.0f / (float)cu);
            const float* r = rows + iv * 12;
            fd.calc(mk3(r[0], r[1], r[2]), mk3(r[3], r[4], r[5]), mk3(r[6], r[7], r[8]), mk3(r[9], r[10], r[11]));
            for (int iu = 0; iu <= cu; ++iu) {
                f3d q = fd.get();
                V[iv * nu + iu] = make_float4(q.x, q.y, q.z, 0.0f);
            }
        }
    } else {
        // surfaceNaturalDice + BilinearEvaluate (surface.cpp:365-388, bilinear.h:189-225)
        const float* c = prm;
        for (int i = tid; i < nv; i += blockDim.x) {
            int iu = i % nu, iv = i / nu;
            float v = (1.0f / (float)cv) * iv;
            float u = (1.0f / (float)cu) * iu;
            f3d A = mk3(c[0], c[1], c[2]), B = mk3(c[3], c[4], c[5]), C = mk3(c[6], c[7], c[8]),
                D = mk3(c[9], c[10], c[11]);
            f3d AB, CD, R;
            if (u <= 0.0f) { AB = A; CD = C; }
            else if (u >= 1.0f) { AB = B; CD = D; }
            else {
                AB = add3(mul3(mk3(B.x - A.x, B.y - A.y, B.z - A.z), u), A);
                CD = add3(mul3(mk3(D.x - C.x, D.y - C.y, D.z - C.z), u), C);
            }
            if (v <= 0.0f) R = AB;
            else if (v >= 1.0f) R = CD;
            else R = add3(mul3(mk3(CD.x - AB.x, CD.y - AB.y, CD.z - AB.z), v), AB);
            V[i] = make_float4(R.x, R.y, R.z, 0.0f);
        }
    }
    __syncthreads();

}

// Floats of one key's parameters (the layouts the host schedule writes).
__device__ __forceinline__ int key_param_floats(int type, int cu, int cv) {
    return type == RH_PRIM_SPHERE ? SPH_TABLES_D + 2 * (cu + 1) + 2 * (cv + 1)
                                  : (type == RH_PRIM_BICUBIC ? 48 : 12);
}

__global__ void __launch_bounds__(256) rh_dice_bust(const FrameDev* __restrict__ fr, const GridDev* __restrict__ grids,
                                                    const PrimDev* __restrict__ prims,
                                                    const float* __restrict__ params,
                                                    float* __restrict__ mprec, float4* __restrict__ mpbound,
                                                    float4* __restrict__ mprecMB) {
    __shared__ float4 V[RH_MAX_GRID_VERTS];
    __shared__ float4 V1[RH_MAX_GRID_VERTS];
    __shared__ float rows[4 * RH_MAX_GRID_ROWS * 3];
    const int tid = threadIdx.x;
    const GridDev G = grids[blockIdx.x];
    const int cu = G.cu, cv = G.cv, nu = cu + 1, nv = (cu + 1) * (cv + 1);
    const float* prm = params + G.paramOff;

    const int type = G.type & ~RH_GRID_MOTION;
    const bool motion = (G.type & RH_GRID_MOTION) != 0;
    dice_points(type, cu, cv, prm, V, rows);
    __syncthreads();
    // Projection to hybrid raster space keeping camera z (micropolygon.cpp:723-731)
    for (int i = tid; i < nv; i += blockDim.x) {
        float4 p = V[i];
        f3d r = xform_point(fr->camToRaster, false, mk3(p.x, p.y, p.z));
        V[i] = make_float4(r.x, r.y, p.z, 0.0f);
    }
    // A deforming surface's grid: key 1 diced with key 0's sizes and
    // projected the same way (CqDeformingSurface::Dice, surface.h:643-655;
    // CqMotionMicroPolyGrid::Split, micropolygon.cpp:970-1024).
    const float* prm1 = prm + key_param_floats(type, cu, cv);
    if (motion) {
        __syncthreads();
        dice_points(type, cu, cv, prm1, V1, rows);
        __syncthreads();
        for (int i = tid; i < nv; i += blockDim.x) {
            float4 p = V1[i];
            f3d r = xform_point(fr->camToRaster, false, mk3(p.x, p.y, p.z));
            V1[i] = make_float4(r.x, r.y, p.z, 0.0f);
        }
    }
    __syncthreads();

    // One hit-test record per micropolygon.
    const int nmp = cu * cv;
    for (int m = tid; m < nmp; m += blockDim.x) {
        const int iu = m % cu, iv = m / cu;
        const int idx = iv * nu + iu;
        float4 q0 = V[idx], q1 = V[idx + 1], q2 = V[idx + nu], q3 = V[idx + nu + 1];
        f3d P[4] = {mk3(q0.x, q0.y, q0.z), mk3(q1.x, q1.y, q1.z), mk3(q2.x, q2.y, q2.z), mk3(q3.x, q3.y, q3.z)};
        // ComputeVertexOrder (micropolygon.cpp:1207-1284): A,B,C,D = P0,P1,P3,P2
        f3d pA = P[0], pB = P[1], pC = P[3], pD = P[2];
        int codeA = 0, codeB = 1, codeC = 3, codeD = 2;
        auto m2 = [](f3d a, f3d b) {
            float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
            return dx * dx + dy * dy + dz * dz;
        };
        f3d vA = pA, vB = pB, vC = pC;
        if ((double)m2(pA, pB) < 1e-8) {
            vB = pC; codeB = codeC; vC = pD; codeC = codeD; codeD = -1;
        } else if ((double)m2(pB, pC) < 1e-8) {
            vB = pC; codeB = codeC; vC = pD; codeC = codeD; codeD = -1;
        } else if ((double)m2(pC, pD) < 1e-8) {
            vC = pD; codeC = codeD; codeD = -1;
        } else if ((double)m2(pD, pA) < 1e-8) {
            codeD = -1;
        }
        bool flip = ((vA.x - vB.x) * (vB.y - vC.y)) >= ((vA.y - vB.y) * (vB.x - vC.x));
        int code;
        bool degen = codeD == -1;
        if (!flip)
            code = degen ? ((codeA & 3) | ((codeC & 3) << 2) | ((codeB & 3) << 4))
                         : ((codeA & 3) | ((codeD & 3) << 2) | ((codeC & 3) << 4) | ((codeB & 3) << 6));
        else
            code = degen ? ((codeA & 3) | ((codeB & 3) << 2) | ((codeC & 3) << 4))
                         : ((codeA & 3) | ((codeB & 3) << 2) | ((codeC & 3) << 4) | ((codeD & 3) << 6));
        // CqInvBilinear::setVertices (bilinear.h:260-275)
        float Ex = P[1].x - P[0].x, Ey = P[1].y - P[0].y;
        float Fx = P[2].x - P[0].x, Fy = P[2].y - P[0].y;
        float Gx = ((-Ex) - P[2].x) + P[3].x, Gy = ((-Ey) - P[2].y) + P[3].y;
        float patchSize = fmax_ref(fmax_ref(fabsf(Fx), fabsf(Fy)), fmax_ref(fabsf(Ex), fabsf(Ey)));
        float irregularity = fmax_ref(fabsf(Gx), fabsf(Gy));
        bool linear = (double)irregularity < 1e-2 * (double)patchSize;
        // CalculateBound (micropolygon.cpp:1663-1673): min(min(min(A,B),C),D)
        float bminx = fmin_ref(fmin_ref(fmin_ref(pA.x, pB.x), pC.x), pD.x);
        float bminy = fmin_ref(fmin_ref(fmin_ref(pA.y, pB.y), pC.y), pD.y);
        float bminz = fmin_ref(fmin_ref(fmin_ref(pA.z, pB.z), pC.z), pD.z);
        float bmaxx = fmax_ref(fmax_ref(fmax_ref(pA.x, pB.x), pC.x), pD.x);
        float bmaxy = fmax_ref(fmax_ref(fmax_ref(pA.y, pB.y), pC.y), pD.y);
        // isOpaque = Oi >= (1,1,1) component-wise (micropolygon.cpp:1482,
        // color.h:432-435); a matte alpha surface's hits are stored as
        // occluding hits as an opaque one's (bucketprocessor.cpp:1490-1491)
        const PrimDev& Pr = prims[G.prim];
        const bool opaque = (Pr.Oi[0] >= 1.0f && Pr.Oi[1] >= 1.0f && Pr.Oi[2] >= 1.0f) || Pr.matte == RH_DEV_MATTE_ALPHA;
        float4* out = reinterpret_cast<float4*>(mprec + (size_t)(G.mpBase + m) * RH_MPREC_FLOATS);
        const uint32_t flags = (uint32_t)(code & 0xff) | (degen ? 1u << 8 : 0u) | (linear ? 1u << 9 : 0u) |
                               (opaque ? 1u << 10 : 0u);
        out[0] = make_float4(P[0].x, P[0].y, P[1].x, P[1].y);
        out[1] = make_float4(P[2].x, P[2].y, P[3].x, P[3].y);
        out[2] = make_float4(P[0].z, P[1].z, P[2].z, P[3].z);
        out[3] = make_float4(bminx, bmaxx, bminy, bmaxy);
        out[4] = make_float4(bminz, __uint_as_float(flags), __int_as_float((int)blockIdx.x), 0.0f);
        if (!mprecMB) {
            if (fr->dofOn) {
                // AddMPG grows the bound by the larger circle of confusion at
                // its two depths (imagebuffer.cpp:519-530); binning reads it
                // for the bucket test and the pixel window.
                const float bmaxz = fmax_ref(fmax_ref(fmax_ref(pA.z, pB.z), pC.z), pD.z);
                const float c0 = fr->dofMult * fabsf(1.0f / bminz - fr->dofOneOverF);
                const float c1 = fr->dofMult * fabsf(1.0f / bmaxz - fr->dofOneOverF);
                const float cx = fmax_ref(fr->dofScaleX * c0, fr->dofScaleX * c1);
                const float cy = fmax_ref(fr->dofScaleY * c0, fr->dofScaleY * c1);
                mpbound[G.mpBase + m] = make_float4(bminx - cx, bmaxx + cx, bminy - cy, bmaxy + cy);
            } else {
                mpbound[G.mpBase + m] = make_float4(bminx, bmaxx, bminy, bmaxy);
            }
            continue;
        }
        // Motion frames: every micropolygon also gets the record the motion
        // sampler reads: both keys' vertices in grid order (a static
        // micropolygon repeats key 0), flags, grid, key times, and its pixel
        // window bound: the static bound, or for a moving one the hull of its
        // sub-bounds, which reach past the keys when the shutter does.  A
        // moving one meets AddMPG's bucket test with m_Bound, key 0's bound
        // encapsulating key 1's (AppendKey, micropolygon.cpp:1952-1967).
        f3d K1[4] = {P[0], P[1], P[2], P[3]};
        if (motion) {
            float4 r0 = V1[idx], r1 = V1[idx + 1], r2 = V1[idx + nu], r3 = V1[idx + nu + 1];
            K1[0] = mk3(r0.x, r0.y, r0.z); K1[1] = mk3(r1.x, r1.y, r1.z);
            K1[2] = mk3(r2.x, r2.y, r2.z); K1[3] = mk3(r3.x, r3.y, r3.z);
        }
        float4* mb = mprecMB + (size_t)(G.mpBase + m) * (RH_MPREC_MB_FLOATS / 4);
        mb[0] = make_float4(P[0].x, P[0].y, P[0].z, P[1].x);
        mb[1] = make_float4(P[1].y, P[1].z, P[2].x, P[2].y);
        mb[2] = make_float4(P[2].z, P[3].x, P[3].y, P[3].z);
        mb[3] = make_float4(K1[0].x, K1[0].y, K1[0].z, K1[1].x);
        mb[4] = make_float4(K1[1].y, K1[1].z, K1[2].x, K1[2].y);
        mb[5] = make_float4(K1[2].z, K1[3].x, K1[3].y, K1[3].z);
        const float t0 = motion ? prm1[key_param_floats(type, cu, cv)] : 0.0f;
        const float t1 = motion ? prm1[key_param_floats(type, cu, cv) + 1] : 0.0f;
        mb[6] = make_float4(__uint_as_float(flags | (motion ? RH_MB_MOVING : 0u)), __int_as_float((int)blockIdx.x), t0, t1);
        if (motion) {
            float kb[2][6];
            mb_key_bound(P, kb[0]);
            mb_key_bound(K1, kb[1]);
            mpbound[G.mpBase + m] = make_float4(fmin_ref(kb[0][0], kb[1][0]), fmax_ref(kb[0][3], kb[1][3]),
                                                fmin_ref(kb[0][1], kb[1][1]), fmax_ref(kb[0][4], kb[1][4]));
            const int divisions = mb_divisions(kb[0], P[0], K1[0], max(4, fr->nx * fr->ny));
            mb[7] = mb_hull(kb, divisions, t0, t1, fr->shutterOpen, fr->shutterClose);
        } else {
            mpbound[G.mpBase + m] = make_float4(bminx, bmaxx, bminy, bmaxy);
            mb[7] = make_float4(bminx, bmaxx, bminy, bmaxy);
        }
    }
}

// ---------------------------------------------------------------------------
// Binning: the ordered micropolygon list of each sampling task
// ---------------------------------------------------------------------------
// A segment is one (bucket, candidate grid) pair; a bucket's segments are in
// dicing order and a grid has at most 256 micropolygons.  Every micropolygon goes through
// CqImageBuffer::AddMPG's tests for the bucket (imagebuffer.cpp:514-591),
// and a kept one belongs to the list of each of the bucket's tasks that its
// clipped pixel window meets.  Three passes keep the reference's order with
// the work spread evenly however uneven the buckets are: count per segment
// and task, an exclusive scan over each bucket's segments, then an ordered
// write.  Lists never exceed the host's capacities; an overflow is counted
// in counters[5].
#define RH_MAX_TASKS 16
#define RH_SEG_STRIDE (RH_MAX_TASKS + 1)   // per segment: a count per task, then kept

// The frame's constants of AddMPG's bucket test, read once per wave.
struct BinConst {
    float dx0, dy0, dx1, dy1;   // crop window grown by half the filter width
    float dsx, dsy;             // bound growth: lfloor(filter width / 2)
    float bxs, bys;             // bucket size
    float ibx, iby;             // 1 / bucket size (used when a power of two)
};

__device__ __forceinline__ BinConst bin_const(const FrameDev* __restrict__ fr) {
    BinConst k;
    const float hfx = fr->filterX * 0.5f, hfy = fr->filterY * 0.5f;   // xFiltSize / 2.0f
    k.dx0 = fr->cropX0 - hfx; k.dy0 = fr->cropY0 - hfy;
    k.dx1 = fr->cropX1 + hfx; k.dy1 = fr->cropY1 + hfy;
    k.dsx = (float)fr->shiftX; k.dsy = (float)fr->shiftY;
    k.bxs = (float)fr->bucketX; k.bys = (float)fr->bucketY;
    k.ibx = 1.0f / k.bxs; k.iby = 1.0f / k.bys;
    return k;
}

// AddMPG for bucket B (imagebuffer.cpp:514-591) with bound bd, then the
// pixel range of window bound wd clipped to the bucket's sample region
// (RenderMPG_Static, bucketprocessor.cpp:1120-1140; for a moving
// micropolygon wd is the hull of the sub-bounds RenderMPG_MBOrDof visits,
// :1290-1310), relative to the region; empty ranges are not kept.
// Branch-free.  The bucket index clamps of AddMPG are left out:
// B lies inside the bucket region, so they never change the outcome.  A
// division by a power-of-two bucket size equals the multiplication by its
// reciprocal (both round the same exact quotient).
template <bool POW2>
__device__ __forceinline__ bool bin_window(const BinConst& k, const BucketDev& B, const float4 bd, const float4 wd,
                                           int& sX, int& eX, int& sY, int& eY) {
    const bool discard = (bd.y < k.dx0) | (bd.w < k.dy0) | (bd.x > k.dx1) | (bd.z > k.dy1);
    const float ex0 = bd.x - k.dsx, ey0 = bd.z - k.dsy, ex1 = bd.y + k.dsx, ey1 = bd.w + k.dsy;
    const int iXBa = (int)(POW2 ? ex0 * k.ibx : ex0 / k.bxs), iYBa = (int)(POW2 ? ey0 * k.iby : ey0 / k.bys);
    const int iXBb = (int)(POW2 ? ex1 * k.ibx : ex1 / k.bxs), iYBb = (int)(POW2 ? ey1 * k.iby : ey1 / k.bys);
    const bool inB = (iXBa <= B.bx) & (B.bx <= iXBb) & (iYBa <= B.by) & (B.by <= iYBb);
    eX = min(lceil_i(wd.y), B.sx1);
    eY = min(lceil_i(wd.w), B.sy1);
    sX = max(floor_i(wd.x), B.sx0);
    sY = max(floor_i(wd.z), B.sy0);
    const bool keep = !discard & inB & (sX < eX) & (sY < eY);
    sX -= B.sx0; eX -= B.sx0; sY -= B.sy0; eY -= B.sy0;
    return keep;
}

// A bucket's tasks form a row-major grid of B.ntx x B.nty rectangles; the
// column and row ranges relative to the bucket's sample region (columns or
// rows beyond the grid are empty ranges that nothing meets).
template <int MT>
__device__ __forceinline__ void task_ranges(const TaskDev* __restrict__ tasks, const BucketDev& B, int* cLo,
                                            int* cHi, int* rLo, int* rHi) {
#pragma unroll
    for (int c = 0; c < MT; ++c) {
        cLo[c] = INT_MAX;
        cHi[c] = INT_MIN;
        if (c < B.ntx) {
            const TaskDev T = tasks[B.task0 + c];
            cLo[c] = T.x0 - B.sx0;
            cHi[c] = T.x1 - B.sx0;
        }
    }
#pragma unroll
    for (int r = 0; r < MT; ++r) {
        rLo[r] = INT_MAX;
        rHi[r] = INT_MIN;
        if (r < B.nty) {
            const TaskDev T = tasks[B.task0 + r * B.ntx];
            rLo[r] = T.y0 - B.sy0;
            rHi[r] = T.y1 - B.sy0;
        }
    }
}

// Pass 1: per segment, the kept micropolygons per task and in all.  One
// wave per segment at a time (grid-stride over segments), 64 micropolygons
// per step; no barriers.  MT: the largest task grid side (2 or 4).
template <int MT, bool POW2, bool MB>
__global__ void __launch_bounds__(256) rh_bin_count(const FrameDev* __restrict__ fr,
                                                    const BucketDev* __restrict__ buckets,
                                                    const TaskDev* __restrict__ tasks, const GridDev* __restrict__ grids,
                                                    const uint32_t* __restrict__ candGrid,
                                                    const uint32_t* __restrict__ segBucket, size_t nseg,
                                                    const float4* __restrict__ mpbound,
                                                    const float4* __restrict__ mprecMB,
                                                    uint32_t* __restrict__ segCounts) {
    const int lane = threadIdx.x & 63;
    const BinConst K = bin_const(fr);
    const size_t waves = (size_t)gridDim.x * (blockDim.x >> 6);
    for (size_t seg = (size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); seg < nseg; seg += waves) {
        const BucketDev B = buckets[segBucket[seg]];
        const GridDev G = grids[candGrid[seg]];
        const int nmp = G.cu * G.cv;
        int cLo[MT], cHi[MT], rLo[MT], rHi[MT];
        task_ranges<MT>(tasks, B, cLo, cHi, rLo, rHi);
        uint32_t cnt[MT * MT], all = 0;
#pragma unroll
        for (int t = 0; t < MT * MT; ++t) cnt[t] = 0;
        for (int base = 0; base < nmp; base += 64) {
            const int m = base + lane;
            const uint32_t id = G.mpBase + min(m, nmp - 1);
            const float4 bd = mpbound[id];
            const float4 wd = MB ? mprecMB[(size_t)id * (RH_MPREC_MB_FLOATS / 4) + 7] : bd;
            int sX, eX, sY, eY;
            const bool keep = (m < nmp) & bin_window<POW2>(K, B, bd, wd, sX, eX, sY, eY);
            all += (uint32_t)__popcll(__ballot(keep));
            bool co[MT], ro[MT];
#pragma unroll
            for (int c = 0; c < MT; ++c) co[c] = keep & (sX < cHi[c]) & (eX > cLo[c]);
#pragma unroll
            for (int r = 0; r < MT; ++r) ro[r] = (sY < rHi[r]) & (eY > rLo[r]);
#pragma unroll
            for (int r = 0; r < MT; ++r)
#pragma unroll
                for (int c = 0; c < MT; ++c) cnt[r * MT + c] += (uint32_t)__popcll(__ballot(co[c] & ro[r]));
        }
        if (lane < B.ntask) {
            const int r = lane / B.ntx, c = lane - r * B.ntx;
            uint32_t v = 0;
#pragma unroll
            for (int t = 0; t < MT * MT; ++t) v = t == r * MT + c ? cnt[t] : v;
            segCounts[seg * RH_SEG_STRIDE + lane] = v;
        }
        if (lane == 0) segCounts[seg * RH_SEG_STRIDE + RH_MAX_TASKS] = all;
    }
}

// Pass 2: per bucket, the exclusive scan of each task's counts over the
// bucket's segments (in place), list lengths and the bucket's kept count.
__global__ void __launch_bounds__(256) rh_bin_scan(const BucketDev* __restrict__ buckets,
                                                   const uint32_t* __restrict__ candStart,
                                                   const unsigned long long* __restrict__ capStart,
                                                   uint32_t* __restrict__ segCounts, uint32_t* __restrict__ listCount,
                                                   uint32_t* __restrict__ bucketRefs,
                                                   unsigned long long* __restrict__ counters) {
    __shared__ uint32_t wsum[4];
    const int b = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const BucketDev B = buckets[b];
    const uint32_t s0 = candStart[b], s1 = candStart[b + 1];
    for (int t = 0; t <= B.ntask; ++t) {
        const int k = t < B.ntask ? t : RH_MAX_TASKS;
        uint32_t run = 0;
        for (uint32_t c0 = s0; c0 < s1; c0 += 256) {
            const uint32_t s = c0 + tid;
            const uint32_t v = s < s1 ? segCounts[(size_t)s * RH_SEG_STRIDE + k] : 0u;
            // inclusive scan in the wave
            uint32_t x = v;
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t y = __shfl_up(x, off);
                if (lane >= off) x += y;
            }
            if (lane == 63) wsum[wave] = x;
            __syncthreads();
            uint32_t base = run, total = 0;
            for (int w = 0; w < 4; ++w) {
                if (w < wave) base += wsum[w];
                total += wsum[w];
            }
            if (s < s1 && k < RH_MAX_TASKS) segCounts[(size_t)s * RH_SEG_STRIDE + k] = base + x - v;
            run += total;
            __syncthreads();
        }
        if (tid == 0) {
            if (k == RH_MAX_TASKS) {
                bucketRefs[b] = run;
            } else {
                const unsigned long long cap = capStart[B.task0 + t + 1] - capStart[B.task0 + t];
                listCount[B.task0 + t] = (unsigned long long)run < cap ? run : (uint32_t)cap;
                if ((unsigned long long)run > cap) atomicAdd(&counters[5], 1ull);
            }
        }
    }
}

// Pass 3: per segment, each task's kept micropolygons in order at the
// segment's offset.  One wave per segment at a time, as pass 1.
template <int MT, bool POW2, bool MB>
__global__ void __launch_bounds__(256) rh_bin_write(const FrameDev* __restrict__ fr,
                                                    const BucketDev* __restrict__ buckets,
                                                    const TaskDev* __restrict__ tasks, const GridDev* __restrict__ grids,
                                                    const uint32_t* __restrict__ candGrid,
                                                    const uint32_t* __restrict__ segBucket, size_t nseg,
                                                    const float4* __restrict__ mpbound,
                                                    const float4* __restrict__ mprecMB,
                                                    const uint32_t* __restrict__ segCounts,
                                                    const unsigned long long* __restrict__ capStart,
                                                    uint32_t* __restrict__ lists) {
    const int lane = threadIdx.x & 63;
    const BinConst K = bin_const(fr);
    const size_t waves = (size_t)gridDim.x * (blockDim.x >> 6);
    for (size_t seg = (size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); seg < nseg; seg += waves) {
        const BucketDev B = buckets[segBucket[seg]];
        const GridDev G = grids[candGrid[seg]];
        const int nmp = G.cu * G.cv;
        int cLo[MT], cHi[MT], rLo[MT], rHi[MT];
        task_ranges<MT>(tasks, B, cLo, cHi, rLo, rHi);
        // per task of the grid: next position, list start and capacity
        uint32_t pos[MT * MT], lim[MT * MT];
        unsigned long long start[MT * MT];
#pragma unroll
        for (int r = 0; r < MT; ++r)
#pragma unroll
            for (int c = 0; c < MT; ++c) {
                pos[r * MT + c] = 0;
                lim[r * MT + c] = 0;
                start[r * MT + c] = 0;
                if (r < B.nty && c < B.ntx) {
                    const int t = B.task0 + r * B.ntx + c;
                    pos[r * MT + c] = segCounts[seg * RH_SEG_STRIDE + (t - B.task0)];
                    start[r * MT + c] = capStart[t];
                    lim[r * MT + c] = (uint32_t)(capStart[t + 1] - capStart[t]);
                }
            }
        for (int base = 0; base < nmp; base += 64) {
            const int m = base + lane;
            const uint32_t id = G.mpBase + min(m, nmp - 1);
            const float4 bd = mpbound[id];
            const float4 wd = MB ? mprecMB[(size_t)id * (RH_MPREC_MB_FLOATS / 4) + 7] : bd;
            int sX, eX, sY, eY;
            const bool keep = (m < nmp) & bin_window<POW2>(K, B, bd, wd, sX, eX, sY, eY);
            if (!__ballot(keep)) continue;
            bool co[MT], ro[MT];
#pragma unroll
            for (int c = 0; c < MT; ++c) co[c] = keep & (sX < cHi[c]) & (eX > cLo[c]);
#pragma unroll
            for (int r = 0; r < MT; ++r) ro[r] = (sY < rHi[r]) & (eY > rLo[r]);
#pragma unroll
            for (int r = 0; r < MT; ++r)
#pragma unroll
                for (int c = 0; c < MT; ++c) {
                    const int t = r * MT + c;
                    const bool ok = co[c] & ro[r];
                    const unsigned long long msk = __ballot(ok);
                    if (ok) {
                        const uint32_t p = pos[t] + (uint32_t)__popcll(msk & ((1ull << lane) - 1ull));
                        if (p < lim[t]) lists[start[t] + p] = G.mpBase + (uint32_t)m;
                    }
                    pos[t] += (uint32_t)__popcll(msk);
                }
        }
    }
}

// Frame counters from the sampling rows and per-bucket rows: counters[0..4] the
// hider's SPL counts, counters[6] the micropolygon-bucket references.
// One workgroup.
__global__ void __launch_bounds__(1024) rh_reduce_counters(const uint32_t* __restrict__ rowCounts, int nrows,
                                                           const uint32_t* __restrict__ bucketRefs, int nbuckets,
                                                           unsigned long long* __restrict__ counters) {
    __shared__ unsigned long long part[6][32];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    unsigned long long acc[6] = {0, 0, 0, 0, 0, 0};
    for (int t = tid; t < nrows; t += 1024)
        for (int c = 0; c < 5; ++c) acc[c] += rowCounts[(size_t)t * 5 + c];
    for (int b = tid; b < nbuckets; b += 1024) acc[5] += bucketRefs[b];
    for (int c = 0; c < 6; ++c) {
        unsigned long long v = acc[c];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) part[c][wave] = v;
    }
    __syncthreads();
    if (tid < 6) {
        unsigned long long v = 0;
        for (int w = 0; w < 16; ++w) v += part[tid][w];
        counters[tid < 5 ? tid : 6] = v;
    }
}

// ---------------------------------------------------------------------------
// Sampling: per-sample visibility in the reference's micropolygon order
// ---------------------------------------------------------------------------
// One 256-thread workgroup per work item (blocks of a task, a part of a
// bucket's owned sample region).  The task's ordered micropolygon list is
// staged through LDS in chunks of CH hit-test records.  Each wave owns up to
// SLOTS blocks of 64 samples; lane = one sample, in row-major order of the
// block's sub-sample grid, so a candidate rectangle is a 64-bit lane mask.
// For every chunk and block the wave compacts the micropolygons whose
// candidate rectangle meets the block, leaving out (but counting as tested)
// those whose minimum depth lies beyond every sample's current opaque depth,
// which the occlusion test would skip anyway; then, 64 at a time,
//  (a) for each of them in turn, the lanes of its rectangle test their sample
//      against its bound (order-free, uniform across the wave), keeping a
//      bit mask of the bound hits;
//  (b) each lane pops its bound hits in list order and runs the
//      order-dependent part: the occlusion skip against its current depth,
//      point-in-polygon, inverse bilinear, depth and the store test.
// The candidate rectangle is the exact set of samples RenderMPG_Static visits
// (bucketprocessor.cpp:1120-1216): the pixel range [floor(bminx),
// lceil(bmaxx)) clipped to the bucket's sample region with the first/last
// pixel sub-sample clip, in sub-sample grid units, so every sample sees the
// reference's sequence of tests.
// The sampling form of a micropolygon record, from its compact form (the
// four raster vertices in grid order, their depths, bound, minimum depth,
// flags and grid): cachePointInPolyTest's edge equations on the vertices in
// the order the vertex-order code gives (micropolygon.cpp:1346-1392, a
// degenerate micropolygon repeats its closing edge) and CqInvBilinear's
// set-up on the grid-order vertices (bilinear.h:260-275).  Out: YM, XM, X, Y,
// z, (A, E), (F, G), bound, (minz, linear, grid, opaque).
// Lanes of a wave block are samples in row-major order of the block's
// sub-sample grid (W = blockW * nx columns, H = blockH * ny rows, W * H <= 64).
// below(n): the lanes 0..n-1.
__device__ __forceinline__ unsigned long long lanes_below(int n) {
    return n >= 64 ? ~0ull : (1ull << n) - 1ull;
}

// The lanes whose samples lie in a candidate rectangle (absolute sub-sample
// grid ranges [lo, hi) packed lo | hi << 16), for the block whose first
// column and row are BX0, BY0.  rep has bit r * W set for every row r.
__device__ __forceinline__ unsigned long long block_rect_lanes(uint2 mt, int BX0, int BY0, int W, int H,
                                                               unsigned long long rep) {
    const int x0 = max((int)(mt.x & 0xffffu) - BX0, 0), x1 = min((int)(mt.x >> 16) - BX0, W);
    const int y0 = max((int)(mt.y & 0xffffu) - BY0, 0), y1 = min((int)(mt.y >> 16) - BY0, H);
    if (x1 <= x0 || y1 <= y0) return 0ull;
    const unsigned long long cols = (lanes_below(x1) & ~lanes_below(x0)) * rep;
    return cols & lanes_below(y1 * W) & ~lanes_below(y0 * W);
}

__device__ __forceinline__ float pick4(int i, float a0, float a1, float a2, float a3) {
    return i == 0 ? a0 : (i == 1 ? a1 : (i == 2 ? a2 : a3));
}

__device__ __forceinline__ void expand_record(const float4* __restrict__ src, uint32_t id, float4* out) {
    const float4 a = src[0], b = src[1], zz = src[2], bd = src[3], r4 = src[4];
    const uint32_t fl = __float_as_uint(r4.y);
    const int code = (int)(fl & 0xffu);
    const bool degen = (fl >> 8) & 1u;
    // the vertices in the order the code gives
    float qx[4], qy[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int o = i < 3 ? (code >> (2 * (i + 1))) & 3 : code & 3;
        qx[i] = pick4(o, a.x, a.z, b.x, b.z);
        qy[i] = pick4(o, a.y, a.w, b.y, b.w);
    }
    float YM[4], XM[4], X[4], Y[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int j = (i + 3) & 3;
        YM[i] = qx[i] - qx[j];
        XM[i] = qy[i] - qy[j];
        X[i] = qx[j];
        Y[i] = qy[j];
    }
    if (degen) {
#pragma unroll
        for (int i = 2; i < 4; ++i) {
            YM[i] = qx[3] - qx[1];
            XM[i] = qy[3] - qy[1];
            X[i] = qx[1];
            Y[i] = qy[1];
        }
    }
    const float Ex = a.z - a.x, Ey = a.w - a.y;
    const float Fx = b.x - a.x, Fy = b.y - a.y;
    const float Gx = ((-Ex) - b.x) + b.z, Gy = ((-Ey) - b.y) + b.w;
    out[0] = make_float4(YM[0], YM[1], YM[2], YM[3]);
    out[1] = make_float4(XM[0], XM[1], XM[2], XM[3]);
    out[2] = make_float4(X[0], X[1], X[2], X[3]);
    out[3] = make_float4(Y[0], Y[1], Y[2], Y[3]);
    out[4] = zz;
    out[5] = make_float4(a.x, a.y, Ex, Ey);
    out[6] = make_float4(Fx, Fy, Gx, Gy);
    out[7] = bd;
    out[8] = make_float4(r4.x, __uint_as_float((fl >> 9) & 3u), r4.z, __uint_as_float(id));
}

// Colour and opacity of a hit: InterpolateOutputs (micropolygon.cpp:
// 1443-1462) of four equal vertex values, smooth (bilinear in the hit's uv,
// the default) or constant (the first vertex's).  A = (Ci.rgb, depth),
// B = (Oi.rgb, micropolygon id bits).
__device__ __forceinline__ void shade_hit(const PrimDev& P, float u, float v, float D, uint32_t id, float4& A,
                                          float4& B) {
    if (P.constantShading) {
        A = make_float4(P.Ci[0], P.Ci[1], P.Ci[2], D);
        B = make_float4(P.Oi[0], P.Oi[1], P.Oi[2], __uint_as_float(id));
        return;
    }
    const float w0 = (1 - u) * (1 - v), w1 = u * (1 - v), w2 = (1 - u) * v, w3 = u * v;
    A.x = ((w0 * P.Ci[0] + w1 * P.Ci[0]) + w2 * P.Ci[0]) + w3 * P.Ci[0];
    A.y = ((w0 * P.Ci[1] + w1 * P.Ci[1]) + w2 * P.Ci[1]) + w3 * P.Ci[1];
    A.z = ((w0 * P.Ci[2] + w1 * P.Ci[2]) + w2 * P.Ci[2]) + w3 * P.Ci[2];
    A.w = D;
    B.x = ((w0 * P.Oi[0] + w1 * P.Oi[0]) + w2 * P.Oi[0]) + w3 * P.Oi[0];
    B.y = ((w0 * P.Oi[1] + w1 * P.Oi[1]) + w2 * P.Oi[1]) + w3 * P.Oi[1];
    B.z = ((w0 * P.Oi[2] + w1 * P.Oi[2]) + w2 * P.Oi[2]) + w3 * P.Oi[2];
    B.w = __uint_as_float(id);
}

// A sample's occluding hit as a sampler writes it (A, B from shade_hit):
// without a hit list, Combine's no-list branch (imagepixel.cpp:301-323): a
// matte surface's hit is transparent black and the midpoint filter averages
// its depth with the occluding depth; an opaque-depth pass writes the
// occluding depth (the occlusion tree's).
__device__ __forceinline__ void finish_occluding(const FrameDev* __restrict__ fr, const PrimDev& P, bool listed,
                                                 int depthOnly, float occlZ, float4& A, float4& B) {
    if (depthOnly) {
        A.w = occlZ;
        return;
    }
    if (listed) return;
    if (P.matte == RH_DEV_MATTE) {
        A.x = A.y = A.z = 0.0f;
        B.x = B.y = B.z = 0.0f;
    }
    if (fr->depthFilter == RH_DEV_DEPTH_MIDPOINT) A.w = (float)(0.5 * (double)(A.w + occlZ));
}

// The part of a task's list (ascending micropolygon ids, the busting order)
// below idCut: the micropolygons rendered in a bucket before a surface is
// taken from its heap (0xffffffff: the whole list).
__device__ __forceinline__ uint32_t list_cut(const uint32_t* __restrict__ L, uint32_t n, uint32_t idCut) {
    if (idCut == 0xffffffffu) return n;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (L[mid] < idCut) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

#define RH_S_WAVES 4
#define RH_S_GROUP 16
// Waves per SIMD of the 96-record, 2-blocks-per-wave variant: its 95
// registers and 21 KB of LDS leave room for 5 (the others: 4; 6 would spill).
#ifndef RH_WPE_96_2
#define RH_WPE_96_2 5
#endif

// Work items: (task, first block | block count << 16).  CH records per LDS
// chunk, SLOTS blocks of 64 samples per wave per pass.  Three launch forms:
// a static grid over the items (one workgroup each, runs of RH_S_GROUP
// items per XCD when the grid is a multiple of 8 * RH_S_GROUP), the same per
// bucket (itemBase/nStatic select the bucket's items), and persistent
// workgroups pulling items from a queue (persistent != 0; the count is in
// *nDevice).  Each workgroup writes one row of counters at rowBase +
// blockIdx.x.  STORE selects StoreSample's mode (RH_STORE_*): the usual
// occluding hit, the midpoint store, or every hit listed (not cullable).
template <int CH, int SLOTS, int STORE = RH_STORE_OCCLUDE, int WPE = 4>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
rh_sample(const FrameDev* __restrict__ fr, const BucketDev* __restrict__ buckets, const TaskDev* __restrict__ tasks,
          const uint2* __restrict__ items, int itemBase, int nStatic, const uint32_t* __restrict__ nDevice,
          int persistent, uint32_t* __restrict__ queueHead, const unsigned long long* __restrict__ capStart,
          const uint32_t* __restrict__ listCount, const uint32_t* __restrict__ lists,
          const float4* __restrict__ mprec4, const GridDev* __restrict__ grids, const PrimDev* __restrict__ prims,
          const float* __restrict__ pos2d, const float* __restrict__ pos1d, const uchar4* __restrict__ pixPattern,
          float4* __restrict__ samples, float4* __restrict__ hitData, uint32_t* __restrict__ hitNext,
          uint32_t* __restrict__ hitCounter, uint4* __restrict__ combineList, float* __restrict__ frontDepth,
          uint32_t* __restrict__ rowCounts, int rowBase, uint32_t idCut, const uint32_t* __restrict__ itemCut,
          int depthOnly) {
    __shared__ float4 recs[CH * 9];
    __shared__ uint32_t wsum[RH_S_WAVES][5];
    __shared__ uint2 meta[CH];
    __shared__ uint8_t matched[RH_S_WAVES][CH];
    __shared__ unsigned long long batM[RH_S_WAVES][64];
    __shared__ float4 batB[RH_S_WAVES][64];
    __shared__ int sItem;
    // the wave index is uniform: block and slot geometry stay in scalar registers
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ns = fr->ns, nx = fr->nx, ny = fr->ny, bw = fr->blockW, bh = fr->blockH;
    // lane -> (column, row) of the block's sub-sample grid -> pixel, sub-sample
    const int BW = bw * nx, BH = bh * ny;
    const int rx = lane % BW, ry = lane / BW;
    const bool laneOk = lane < BW * BH;
    const int lx = rx / nx, sm = rx - lx * nx, ly = ry / ny, sn = ry - ly * ny;
    const int pix = ly * bw + lx, si = sn * nx + sm;
    unsigned long long rowRep = 0ull;
    for (int r = 0; r < BH; ++r) rowRep |= 1ull << (r * BW);
    uint32_t nTest = 0;
    // wave-uniform counts (ballots): nTest's uniform part and the others
    uint32_t nTestW = 0, nBound = 0, nHit = 0, nIn = 0, nIn2 = 0;
    uint32_t wbBase = 0;                  // the wave's current arena block
    int wbLeft = 0;
    const int itemEnd = persistent ? (int)nDevice[0] : itemBase + nStatic;
    // isCullable (bucketprocessor.cpp:1076-1081): occlusion tests and the
    // occluding hit apply; else every hit is listed
    constexpr bool cullH = STORE != RH_STORE_LIST_ALL;
    constexpr bool MID = STORE == RH_STORE_MIDPOINT;

    for (int round = 0;; ++round) {
    int it;
    if (persistent) {
        if (tid == 0) sItem = (int)atomicAdd(queueHead, 1u);
        __syncthreads();
        it = sItem;
        __syncthreads();
    } else {
        if (round > 0) break;
        if (gridDim.x % (8 * RH_S_GROUP) == 0) {
            // Workgroups are dealt round-robin to the 8 XCDs; runs of
            // RH_S_GROUP consecutive items (neighbours sharing micropolygons)
            // go to one XCD.
            const int kx = blockIdx.x >> 3;
            it = itemBase + ((kx / RH_S_GROUP) * 8 + (blockIdx.x & 7)) * RH_S_GROUP + kx % RH_S_GROUP;
        } else {
            it = itemBase + blockIdx.x;
        }
    }
    if (it >= itemEnd) break;
    const uint2 item = items[it];
    const int task = (int)item.x, blk0 = (int)(item.y & 0xffffu), blkN = (int)(item.y >> 16);
    const TaskDev T = tasks[task];
    const BucketDev B = buckets[T.bucket];
    const int tw = T.x1 - T.x0;
    const int nbx = (tw + bw - 1) / bw;
    const uint32_t* L = lists + capStart[task];
    const uint32_t n = list_cut(L, listCount[task], itemCut ? itemCut[it] : idCut);

    for (int p0 = 0; p0 < blkN; p0 += RH_S_WAVES * SLOTS) {
        bool has[SLOTS], live[SLOTS], lodOk[SLOTS];
        int bxs[SLOTS], bys[SLOTS], bgx[SLOTS], bgy[SLOTS];
        float sx[SLOTS], sy[SLOTS];
        float occlZ[SLOTS], wu[SLOTS], wv[SLOTS];
        float sD[SLOTS];                  // MID: the occluding hit's depth
        uint32_t win[SLOTS], head[SLOTS];
        int nh[SLOTS];
#pragma unroll
        for (int k = 0; k < SLOTS; ++k) {
            head[k] = 0xffffffffu;
            nh[k] = 0;
            sD[k] = FLT_MAX;
            const int blk = blk0 + p0 + wave + RH_S_WAVES * k;
            has[k] = p0 + wave + RH_S_WAVES * k < blkN;
            const int bx = T.x0 + (blk % nbx) * bw, by = T.y0 + (blk / nbx) * bh;
            const int px = bx + lx, py = by + ly;
            live[k] = has[k] && laneOk && px < T.x1 && py < T.y1;
            bxs[k] = bx;
            bys[k] = by;
            bgx[k] = (bx - B.sx0) * nx;
            bgy[k] = (by - B.sy0) * ny;
            occlZ[k] = FLT_MAX;
            wu[k] = 0.f;
            wv[k] = 0.f;
            win[k] = 0xffffffffu;
            sx[k] = 0.f;
            sy[k] = 0.f;
            lodOk[k] = false;
            if (live[k]) {
                const uchar4 pat = pixPattern[(size_t)(py - fr->extY0) * fr->extW + (px - fr->extX0)];
                // position = (x, y) + pattern position (imagepixel.cpp:353)
                sx[k] = (float)px + pos2d[2 * ((size_t)pat.x * ns + si)];
                sy[k] = (float)py + pos2d[2 * ((size_t)pat.x * ns + si) + 1];
                const float lod = pos1d[(size_t)pat.z * ns + si];
                // LevelOfDetailBounds [0, 1] (bucketprocessor.cpp:1183-1189)
                lodOk[k] = !(0.0f > lod || lod >= 1.0f);
            }
        }
        for (uint32_t c0 = 0; c0 < n; c0 += CH) {
            const int cnt = (int)((n - c0) < (uint32_t)CH ? (n - c0) : (uint32_t)CH);
            __syncthreads();
            if (tid < cnt) {
                const uint32_t id = L[c0 + tid];
                float4 r[9];
                expand_record(mprec4 + (size_t)id * (RH_MPREC_FLOATS / 4), id, r);
#pragma unroll
                for (int q = 0; q < 9; ++q) recs[tid * 9 + q] = r[q];
                const float bminx = r[7].x, bmaxx = r[7].y, bminy = r[7].z, bmaxy = r[7].w;
                int eX = lceil_i(bmaxx), eY = lceil_i(bmaxy);
                if (eX > B.sx1) eX = B.sx1;
                if (eY > B.sy1) eY = B.sy1;
                int sX = floor_i(bminx), sY = floor_i(bminy);
                if (sY < B.sy0) sY = B.sy0;
                if (sX < B.sx0) sX = B.sx0;
                uint32_t gx0 = 0, gx1 = 0, gy0 = 0, gy1 = 0;
                if (sX < eX && sY < eY) {
                    const int im = (bminx < sX) ? 0 : floor_i((bminx - sX) * nx);
                    const int in = (bminy < sY) ? 0 : floor_i((bminy - sY) * ny);
                    const int em = (bmaxx > eX) ? nx : lceil_i((bmaxx - (eX - 1)) * nx);
                    const int en = (bmaxy > eY) ? ny : lceil_i((bmaxy - (eY - 1)) * ny);
                    gx0 = (uint32_t)((sX - B.sx0) * nx + im);
                    gx1 = (uint32_t)((eX - 1 - B.sx0) * nx + em);
                    gy0 = (uint32_t)((sY - B.sy0) * ny + in);
                    gy1 = (uint32_t)((eY - 1 - B.sy0) * ny + en);
                }
                meta[tid] = make_uint2(gx0 | (gx1 << 16), gy0 | (gy1 << 16));
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < SLOTS; ++k) {
                if (!has[k]) continue;  // uniform across the wave
                const int BX0 = bgx[k], BX1 = bgx[k] + BW, BY0 = bgy[k], BY1 = bgy[k] + BH;
                const unsigned long long liveM = __ballot(live[k]), inM = __ballot(live[k] && lodOk[k]);
                // The block's farthest opaque depth so far: a micropolygon whose
                // minimum depth lies beyond it is skipped by every sample's
                // occlusion test (depths only decrease), so it is counted as
                // tested and not visited.
                float zmax = live[k] ? occlZ[k] : -FLT_MAX;
                for (int off = 32; off > 0; off >>= 1) zmax = fmaxf(zmax, __shfl_xor(zmax, off));
                // Ordered list of the chunk's micropolygons that meet this block.
                int nm = 0;
                for (int j0 = 0; j0 < cnt; j0 += 64) {
                    const int j = j0 + lane;
                    bool ok = false, hidden = false;
                    uint2 mt = make_uint2(0u, 0u);
                    if (j < cnt) {
                        mt = meta[j];
                        ok = (int)(mt.x & 0xffffu) < BX1 && (int)(mt.x >> 16) > BX0 && (int)(mt.y & 0xffffu) < BY1 &&
                             (int)(mt.y >> 16) > BY0;
                        if (ok && cullH && recs[j * 9 + 8].x > zmax) {
                            hidden = true;
                            ok = false;
                        }
                    }
                    if (hidden) nTest += (uint32_t)__popcll(block_rect_lanes(mt, BX0, BY0, BW, BH, rowRep) & liveM);
                    const unsigned long long msk = __ballot(ok);
                    if (ok) matched[wave][nm + __popcll(msk & ((1ull << lane) - 1ull))] = (uint8_t)j;
                    nm += __popcll(msk);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const float psx = sx[k], psy = sy[k];
                for (int i0 = 0; i0 < nm; i0 += 64) {
                    const int ni = (nm - i0) < 64 ? (nm - i0) : 64;
                    const int ni4 = (ni + 3) & ~3;
                    // The batch: lane i holds matched micropolygon i0 + i as the
                    // lanes of its candidate rectangle that go on to the bound
                    // test, and its bound; it counts the rectangle's live lanes
                    // as tested.  Padding entries have no lanes.
                    if (lane < ni4) {
                        unsigned long long m = 0ull;
                        float4 bd = make_float4(0.f, 0.f, 0.f, 0.f);
                        if (lane < ni) {
                            const int jj = matched[wave][i0 + lane];
                            m = block_rect_lanes(meta[jj], BX0, BY0, BW, BH, rowRep);
                            bd = recs[jj * 9 + 7];
                            nTest += (uint32_t)__popcll(m & liveM);
                        }
                        batM[wave][lane] = m & inM;
                        batB[wave][lane] = bd;
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    // (a) candidate and bound tests, one micropolygon at a time;
                    // the first micropolygon ends in the high bit.
                    uint32_t ph = 0u, pl = 0u;
                    // four entries: broadcast reads first, then branch-free tests
                    auto test4 = [&](int i, uint32_t acc) -> uint32_t {
                        unsigned long long mm[4];
                        float4 bd[4];
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            mm[q] = batM[wave][i + q];
                            bd[q] = batB[wave][i + q];
                        }
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            // CqBound::Contains2D (bound.h:144-151)
                            const bool inB = !(psx < bd[q].x) & !(psx > bd[q].y) & !(psy < bd[q].z) & !(psy > bd[q].w);
                            const bool in = __builtin_amdgcn_inverse_ballot_w64(mm[q]) & inB;
                            acc = acc + acc + (uint32_t)in;
                        }
                        return acc;
                    };
                    const int n0 = ni4 < 32 ? ni4 : 32, n1 = ni4 - n0;
                    for (int i = 0; i < n0; i += 4) ph = test4(i, ph);
                    for (int i = 32; i < ni4; i += 4) pl = test4(i, pl);
                    // bit i of pb: micropolygon i0 + i
                    ph = ph << (32 - n0);
                    pl = n1 > 0 ? pl << (32 - n1) : 0u;
                    unsigned long long pb = ((unsigned long long)__builtin_bitreverse32(pl) << 32) |
                                            (unsigned long long)__builtin_bitreverse32(ph);
                    // (b) the order-dependent tests, each lane its own hits in list order.
                    while (__ballot(pb != 0ull)) {
                        // One hit test per lane; the outcome is applied below in
                        // one place (opaque store or transparent append).
                        bool hit = false, cBound = false, cIn = false, cIn2 = false;
                        float D = 0.f, u = 0.5f, v = 0.5f;
                        uint32_t hflags = 0u, hid = 0u, hgrid = 0u;
                        if (pb != 0ull) {
                            const int i = __ffsll((long long)pb) - 1;
                            pb &= pb - 1ull;
                            const int j = matched[wave][i0 + i];
                            const float4* R = &recs[j * 9];
                            const float4 r8 = R[8];
                            // occlusion skip against the current opaque hit (bucketprocessor.cpp:1179)
                            if (!(cullH && r8.x > occlZ[k])) {
                                cBound = true;
                                // fContains (micropolygon.cpp:1295-1342): edges 0,1 reject on <= 0, edges 2,3 on < 0
                                const float4 YM = R[0], XM = R[1], X = R[2], Y = R[3];
                                const float px_ = sx[k], py_ = sy[k];
                                const float t0 = ((py_ - Y.x) * YM.x) - ((px_ - X.x) * XM.x);
                                const float t1 = ((py_ - Y.y) * YM.y) - ((px_ - X.y) * XM.y);
                                const float t2 = ((py_ - Y.z) * YM.z) - ((px_ - X.z) * XM.z);
                                const float t3 = ((py_ - Y.w) * YM.w) - ((px_ - X.w) * XM.w);
                                if (!(t0 <= 0 || t1 <= 0 || t2 < 0 || t3 < 0)) {
                                    cIn = true;
                                    // CqInvBilinear::operator() (bilinear.h:277-305)
                                    const float4 r5 = R[5], r6 = R[6], Z = R[4];
                                    const float Ax = r5.x, Ay = r5.y, Ex = r5.z, Ey = r5.w, Fx = r6.x, Fy = r6.y,
                                                Gx = r6.z, Gy = r6.w;
                                    {
                                        const float M1x = Ex + Gx * v, M1y = Ey + Gy * v;
                                        const float M2x = Fx + Gx * u, M2y = Fy + Gy * u;
                                        const float bx = (((Ax + Ex * u) + Fx * v) + (Gx * u) * v) - px_;
                                        const float by = (((Ay + Ey * u) + Fy * v) + (Gy * u) * v) - py_;
                                        float det = M1x * M2y - M1y * M2x;
                                        det = 1 / det;
                                        u = u - det * (bx * M2y - by * M2x);
                                        v = v - det * (-(bx * M1y - by * M1x));
                                    }
                                    hflags = __float_as_uint(r8.y);
                                    if ((hflags & 1u) == 0) {
                                        cIn2 = true;
                                        const float M1x = Ex + Gx * v, M1y = Ey + Gy * v;
                                        const float M2x = Fx + Gx * u, M2y = Fy + Gy * u;
                                        const float bx = (((Ax + Ex * u) + Fx * v) + (Gx * u) * v) - px_;
                                        const float by = (((Ay + Ey * u) + Fy * v) + (Gy * u) * v) - py_;
                                        float det = M1x * M2y - M1y * M2x;
                                        if (det != 0) det = 1 / det;
                                        u = u - det * (bx * M2y - by * M2x);
                                        v = v - det * (-(bx * M1y - by * M1x));
                                    }
                                    // bilerp (bilinear.h:228-236)
                                    const float w0 = (1 - v) * (1 - u), w1 = (1 - v) * u, w2 = v * (1 - u),
                                                w3 = v * u;
                                    D = ((w0 * Z.x + w1 * Z.y) + w2 * Z.z) + w3 * Z.w;
                                    // StoreSample (bucketprocessor.cpp:1471-1569): reject when occlZ <= D
                                    hit = !cullH || !(occlZ[k] <= D);
                                    hid = __float_as_uint(r8.w);
                                    hgrid = __float_as_uint(r8.z);
                                }
                            }
                        }
                        nBound += (uint32_t)__popcll(__ballot(cBound));
                        nIn += (uint32_t)__popcll(__ballot(cIn));
                        nIn2 += (uint32_t)__popcll(__ballot(cIn2));
                        nHit += (uint32_t)__popcll(__ballot(hit));
                        // opaque (or matte alpha) with occlusion: the occluding hit
                        const bool opq = hit && (hflags & 2u) != 0 && cullH;
                        if (MID) {
                            // midpoint: the nearest opaque hit is kept and the
                            // second nearest occludes (bucketprocessor.cpp:1502-1525)
                            const float prevZ = win[k] != 0xffffffffu ? sD[k] : FLT_MAX;
                            const bool keep = !(prevZ < D);
                            const bool store = opq && keep;
                            occlZ[k] = opq ? (keep ? prevZ : D) : occlZ[k];
                            win[k] = store ? hid : win[k];
                            wu[k] = store ? u : wu[k];
                            wv[k] = store ? v : wv[k];
                            sD[k] = store ? D : sD[k];
                        } else {
                            occlZ[k] = opq ? D : occlZ[k];
                            win[k] = opq ? hid : win[k];
                            wu[k] = opq ? u : wu[k];
                            wv[k] = opq ? v : wv[k];
                        }
                        // otherwise a new entry of the sample's hit list
                        const bool app = hit && !opq && !depthOnly;
                        // Append the wave's transparent hits (converged): entries come
                        // from 64-entry blocks of the frame's arena, one atomic per block.
                        const unsigned long long am = __ballot(app);
                        if (am) {
                            const int na = __popcll(am);
                            if (na > wbLeft) {
                                uint32_t base = 0;
                                if (lane == 0) base = atomicAdd(hitCounter, 64u);
                                wbBase = __shfl(base, 0);
                                wbLeft = 64;
                            }
                            if (app) {
                                const uint32_t slot =
                                    wbBase + (uint32_t)(64 - wbLeft) + (uint32_t)__popcll(am & ((1ull << lane) - 1ull));
                                if (slot < fr->hitCapacity) {
                                    // the hit's (u, v, depth, id) and surface; the
                                    // combine shades it
                                    hitData[slot] = make_float4(u, v, D, __uint_as_float(hid));
                                    hitNext[2 * (size_t)slot] = head[k];
                                    hitNext[2 * (size_t)slot + 1] = grids[hgrid].prim;
                                    head[k] = slot;
                                    ++nh[k];
                                }
                            }
                            wbLeft -= na;
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < SLOTS; ++k) {
            if (!live[k]) continue;
            const int px = bxs[k] + lx, py = bys[k] + ly;
            const size_t sidx = ((size_t)(py - fr->extY0) * fr->extW + (px - fr->extX0)) * ns + si;
            float4 a = make_float4(0.f, 0.f, 0.f, FLT_MAX), b = make_float4(0.f, 0.f, 0.f, __uint_as_float(win[k]));
            if (win[k] != 0xffffffffu) {
                const float4 r8 = mprec4[(size_t)win[k] * (RH_MPREC_FLOATS / 4) + 4];
                const PrimDev& P = prims[grids[__float_as_int(r8.z)].prim];
                shade_hit(P, wu[k], wv[k], MID ? sD[k] : occlZ[k], win[k], a, b);
                finish_occluding(fr, P, nh[k] > 0, depthOnly, occlZ[k], a, b);
            }
            samples[2 * sidx] = a;
            samples[2 * sidx + 1] = b;
            // the combine's occluding depth for a listed sample, else the
            // sample's final depth
            if (frontDepth) frontDepth[sidx] = nh[k] > 0 ? occlZ[k] : a.w;
            // Samples with transparent hits are combined by rh_combine.
            const bool multi = nh[k] > 0;
            const unsigned long long cm = __ballot(multi);
            if (cm) {
                uint32_t base = 0;
                if (lane == __ffsll((long long)cm) - 1) base = atomicAdd(&hitCounter[2], (uint32_t)__popcll(cm));
                base = __shfl(base, __ffsll((long long)cm) - 1);
                if (multi) {
                    const uint32_t e = base + (uint32_t)__popcll(cm & ((1ull << lane) - 1ull));
                    combineList[e] = make_uint4((uint32_t)sidx, (uint32_t)(sidx >> 32), head[k], (uint32_t)nh[k]);
                }
            }
        }
    }
    }   // items
    if (lane == 0) nTest += nTestW;
    // The workgroup's hider counters: wave sums, then one row (summed by
    // rh_reduce_counters; same-address atomics from every wave would
    // serialise in L2).
    for (int off = 32; off > 0; off >>= 1) nTest += __shfl_xor(nTest, off);
    if (lane == 0) {
        wsum[wave][0] = nTest; wsum[wave][1] = nBound; wsum[wave][2] = nHit; wsum[wave][3] = nIn;
        wsum[wave][4] = nIn2;
    }
    __syncthreads();
    if (tid < 5) {
        uint32_t v = 0;
        for (int w = 0; w < RH_S_WAVES; ++w) v += wsum[w][tid];
        rowCounts[(size_t)(rowBase + blockIdx.x) * 5 + tid] = v;
    }
}

// ---------------------------------------------------------------------------
// Sampling of frames with moving micropolygons
// ---------------------------------------------------------------------------
// The point-in-polygon test of one sample against four raster vertices in
// grid order with a given vertex-order code: cachePointInPolyTest's edge
// equations (micropolygon.cpp:1346-1392), fContains's strictness
// (:1295-1342), CqInvBilinear with one or two Newton steps
// (bilinear.h:260-310) and the depth bilerp (:228-236).
struct MBHit {
    bool in;
    bool linear;
    float D, u, v;
};

__device__ __forceinline__ MBHit quad_hit(const f3d p0, const f3d p1, const f3d p2, const f3d p3, uint32_t code,
                                          bool degen, float px, float py) {
    const f3d pin[4] = {p0, p1, p2, p3};
    MBHit h;
    h.in = false;
    h.D = 0.f; h.u = 0.5f; h.v = 0.5f;
    const float Ex = pin[1].x - pin[0].x, Ey = pin[1].y - pin[0].y;
    const float Fx = pin[2].x - pin[0].x, Fy = pin[2].y - pin[0].y;
    const float Gx = ((-Ex) - pin[2].x) + pin[3].x, Gy = ((-Ey) - pin[2].y) + pin[3].y;
    const float patchSize = fmax_ref(fmax_ref(fabsf(Fx), fabsf(Fy)), fmax_ref(fabsf(Ex), fabsf(Ey)));
    const float irregularity = fmax_ref(fabsf(Gx), fabsf(Gy));
    h.linear = (double)irregularity < 1e-2 * (double)patchSize;
    float qx[4], qy[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int o = i < 3 ? (int)(code >> (2 * (i + 1))) & 3 : (int)code & 3;
        qx[i] = o == 0 ? p0.x : (o == 1 ? p1.x : (o == 2 ? p2.x : p3.x));
        qy[i] = o == 0 ? p0.y : (o == 1 ? p1.y : (o == 2 ? p2.y : p3.y));
    }
    float YM[4], XM[4], X[4], Y[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int j = (i + 3) & 3;
        YM[i] = qx[i] - qx[j];
        XM[i] = qy[i] - qy[j];
        X[i] = qx[j];
        Y[i] = qy[j];
    }
    if (degen) {
#pragma unroll
        for (int i = 2; i < 4; ++i) {
            YM[i] = qx[3] - qx[1];
            XM[i] = qy[3] - qy[1];
            X[i] = qx[1];
            Y[i] = qy[1];
        }
    }
    const float t0 = ((py - Y[0]) * YM[0]) - ((px - X[0]) * XM[0]);
    const float t1 = ((py - Y[1]) * YM[1]) - ((px - X[1]) * XM[1]);
    const float t2 = ((py - Y[2]) * YM[2]) - ((px - X[2]) * XM[2]);
    const float t3 = ((py - Y[3]) * YM[3]) - ((px - X[3]) * XM[3]);
    if (t0 <= 0 || t1 <= 0 || t2 < 0 || t3 < 0) return h;
    h.in = true;
    const float Ax = pin[0].x, Ay = pin[0].y;
    float u = 0.5f, v = 0.5f;
    {
        const float M1x = Ex + Gx * v, M1y = Ey + Gy * v;
        const float M2x = Fx + Gx * u, M2y = Fy + Gy * u;
        const float bx = (((Ax + Ex * u) + Fx * v) + (Gx * u) * v) - px;
        const float by = (((Ay + Ey * u) + Fy * v) + (Gy * u) * v) - py;
        float det = M1x * M2y - M1y * M2x;
        det = 1 / det;
        u = u - det * (bx * M2y - by * M2x);
        v = v - det * (-(bx * M1y - by * M1x));
    }
    if (!h.linear) {
        const float M1x = Ex + Gx * v, M1y = Ey + Gy * v;
        const float M2x = Fx + Gx * u, M2y = Fy + Gy * u;
        const float bx = (((Ax + Ex * u) + Fx * v) + (Gx * u) * v) - px;
        const float by = (((Ay + Ey * u) + Fy * v) + (Gy * u) * v) - py;
        float det = M1x * M2y - M1y * M2x;
        if (det != 0) det = 1 / det;
        u = u - det * (bx * M2y - by * M2x);
        v = v - det * (-(bx * M1y - by * M1x));
    }
    const float w0 = (1 - v) * (1 - u), w1 = (1 - v) * u, w2 = v * (1 - u), w3 = v * u;
    h.D = ((w0 * pin[0].z + w1 * pin[1].z) + w2 * pin[2].z) + w3 * pin[3].z;
    h.u = u;
    h.v = v;
    return h;
}

// The motion form of the sampler.  A frame with moving micropolygons visits
// samples per sub-bound (RenderMPG_MBOrDof, bucketprocessor.cpp:1221-1469):
// a sub-bound of CqMicroPolygonMotion::BuildBoundList
// (micropolygon.cpp:1689-1756) visits the samples of its pixels whose index
// lies in the window its times give (sample i's time lies in the i-th of ns
// equal parts of the shutter), each visit counted as a sample test; then the
// time window, bound, occlusion and LOD tests and CqMicroPolygonMotion::Sample
// (:1768-1914): the tight bound and the vertices at the sample's time, and
// the point-in-polygon test on them.  A static micropolygon in the same frame
// is RenderMPG_Static (:1097-1218).  A window reaching past a pixel's
// samples stops at the last one.
//
// One workgroup of NW waves per task.  The task's list is taken in chunks of
// records staged in LDS; for each chunk:
//  1. every (record, sub-bound) pair is evaluated once for the task, by
//     all threads together, and the pairs whose visits meet the task become
//     entries, kept in pair order (the reference's visiting order): the
//     sub-sample rectangle, the sample index window and the sub-bound's two
//     end times;
//  2. each wave holds one block of samples (lanes in row-major sub-sample
//     order, as rh_sample), gathers the entries meeting its block into a
//     buffer of up to 64, and turns them into one 64-bit mask per lane of the
//     entries that visit the lane's sample; the lanes then take their own
//     entries lowest bit first, so each sample sees its visits in order and a
//     wave runs as many visit steps as its busiest lane has visits.
#ifndef RH_MB_RECS
#define RH_MB_RECS 128     // records staged per chunk (a multiple of 64)
#endif
#ifndef RH_MB_PAIRS
#define RH_MB_PAIRS 1792   // (record, sub-bound) pairs per chunk at most
#endif
#define RH_MB_RING 128     // visit results held per wave

// Transpose of a 64 x 64 bit matrix held one row per lane (bit c of lane r's
// word is element (r, c)): lane c returns column c.  Six rounds exchange
// blocks of 32, 16, ... 1 bits with the lane that many rows away.
__device__ __forceinline__ unsigned long long transpose64(unsigned long long row, int lane) {
    const unsigned long long keep[6] = {0x00000000ffffffffull, 0x0000ffff0000ffffull, 0x00ff00ff00ff00ffull,
                                        0x0f0f0f0f0f0f0f0full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int s = 32 >> k;
        const unsigned long long K = keep[k];
        const uint32_t plo = (uint32_t)__shfl_xor((int)(uint32_t)row, s);
        const uint32_t phi = (uint32_t)__shfl_xor((int)(uint32_t)(row >> 32), s);
        const unsigned long long partner = ((unsigned long long)phi << 32) | plo;
        row = (lane & s) ? (((partner >> s) & K) | (row & ~K)) : ((row & K) | ((partner & K) << s));
    }
    return row;
}

// The position of the k-th (from 0) set bit of m, which has more than k.
__device__ __forceinline__ int select_bit(unsigned long long m, int k) {
    int pos = 0;
    int c = __popc((uint32_t)m);
    if (k >= c) { k -= c; m >>= 32; pos = 32; }
    uint32_t w = (uint32_t)m;
    c = __popc(w & 0xffffu);
    if (k >= c) { k -= c; w >>= 16; pos += 16; }
    c = __popc(w & 0xffu);
    if (k >= c) { k -= c; w >>= 8; pos += 8; }
    c = __popc(w & 0xfu);
    if (k >= c) { k -= c; w >>= 4; pos += 4; }
    c = __popc(w & 0x3u);
    if (k >= c) { k -= c; w >>= 2; pos += 2; }
    return pos + ((k >= (int)(w & 1u)) ? 1 : 0);
}

template <int NW, int STORE>
__global__ void __launch_bounds__(NW * 64) rh_sample_mb(const FrameDev* __restrict__ fr,
                                                        const BucketDev* __restrict__ buckets,
                                                        const TaskDev* __restrict__ tasks, int ntasks,
                                                        const unsigned long long* __restrict__ capStart,
                                                        const uint32_t* __restrict__ listCount,
                                                        const uint32_t* __restrict__ lists,
                                                        const float4* __restrict__ mprecMB,
                                                        const GridDev* __restrict__ grids,
                                                        const PrimDev* __restrict__ prims,
                                                        const float* __restrict__ pos2d, const float* __restrict__ pos1d,
                                                        const uchar4* __restrict__ pixPattern,
                                                        float4* __restrict__ samples, float4* __restrict__ hitData,
                                                        uint32_t* __restrict__ hitNext, uint32_t* __restrict__ hitCounter,
                                                        uint4* __restrict__ combineList, float* __restrict__ frontDepth,
                                                        uint32_t* __restrict__ rowCounts, int rowBase, int taskBase,
                                                        uint32_t idCut, const uint32_t* __restrict__ itemCut,
                                                        const uint32_t* __restrict__ taskList, int depthOnly) {
    __shared__ float4 recs[RH_MB_RECS * 8];
    __shared__ float kbs[RH_MB_RECS][12];          // key bounds, kb[k][c] at [k * 6 + c]
    __shared__ float dts[RH_MB_RECS];
    __shared__ int divs[RH_MB_RECS];
    __shared__ int offs[RH_MB_RECS + 1];
    __shared__ uint32_t ids[RH_MB_RECS];
    // entries: (record | sub-bound << 8 (0xff: static) | index window << 16,
    // sub-sample x range, y range (lo | hi << 16), min z); the bound
    // (min x, max x, min y, max y); the times (time0, time1)
    __shared__ uint4 E0[RH_MB_PAIRS];
    __shared__ float4 E1[RH_MB_PAIRS];
    __shared__ float2 E2[RH_MB_PAIRS];
    __shared__ uint16_t buf[NW][128];
    __shared__ float4 smp[NW][64];                 // the block's samples: x, y, time, LOD test
    __shared__ unsigned long long winLanes[65];
    // per wave: the visit results ring, and per rank its first position and lanes
    __shared__ uint32_t ringCode[NW][RH_MB_RING];
    __shared__ float ringD[NW][RH_MB_RING], ringU[NW][RH_MB_RING], ringV[NW][RH_MB_RING];
    __shared__ int rankStart[NW][65];
    __shared__ unsigned long long rankLanes[NW][64];
    __shared__ uint32_t wtot[NW];
    __shared__ int nfitShared;
    __shared__ uint32_t wsum[NW][5];
    const int slot = taskBase + (int)blockIdx.x;
    const int task = (taskList && slot < ntasks) ? (int)taskList[slot] : slot;
    // StoreSample's mode (RH_STORE_*)
    constexpr bool cullH = STORE != RH_STORE_LIST_ALL, midS = STORE == RH_STORE_MIDPOINT;
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const unsigned long long lanesBelow = (1ull << lane) - 1ull;
    uint32_t nTest = 0, nBound = 0, nHit = 0, nIn = 0, nIn2 = 0;
    uint32_t wbBase = 0;
    int wbLeft = 0;
    if (slot < ntasks) {
    const int ns = fr->ns, nx = fr->nx, ny = fr->ny, bw = fr->blockW, bh = fr->blockH;
    const int BW = bw * nx, BH = bh * ny;
    const int rx = lane % BW, ry = lane / BW;
    const bool laneOk = lane < BW * BH;
    const int lx = rx / nx, sm = rx - lx * nx, ly = ry / ny, sn = ry - ly * ny;
    const int si = sn * nx + sm;
    const TaskDev T = tasks[task];
    const BucketDev B = buckets[T.bucket];
    const int nbx = (T.x1 - T.x0 + bw - 1) / bw, nby = (T.y1 - T.y0 + bh - 1) / bh;
    const uint32_t* L = lists + capStart[task];
    const uint32_t n = list_cut(L, listCount[task], itemCut ? itemCut[slot] : idCut);
    const float opentime = fr->shutterOpen, closetime = fr->shutterClose;
    // isClose (math.h:174-182)
    const float dsh = fabsf(closetime - opentime);
    const bool fastShutter = dsh <= 1.1920929e-06f * fabsf(closetime) || dsh <= 1.1920929e-06f * fabsf(opentime);
    const float timePerSample = fastShutter ? 0.0f : ns / (closetime - opentime);
    const int timeRanges = max(4, nx * ny);
    const float clipNear = fr->clipNear, clipFar = fr->clipFar;
    // the task in sub-sample units of the bucket's sample region
    const int tgx0 = (T.x0 - B.sx0) * nx, tgx1 = (T.x1 - B.sx0) * nx;
    const int tgy0 = (T.y0 - B.sy0) * ny, tgy1 = (T.y1 - B.sy0) * ny;
    // a block's lanes of sub-sample index below k, and the rows of its lanes
    unsigned long long rowRep = 0ull;
    for (int r = 0; r < BH; ++r) rowRep |= 1ull << (r * BW);
    if (wave == 0)
        for (int k = 0; k <= 64; ++k) {
            const unsigned long long m = __ballot(laneOk && si < k);
            if (lane == 0) winLanes[k] = m;
        }

    for (int p0 = 0; p0 < nbx * nby; p0 += NW) {
        const int blk = p0 + wave;
        const bool has = blk < nbx * nby;       // uniform per wave
        const int bx = T.x0 + (blk % nbx) * bw, by = T.y0 + (blk / nbx) * bh;
        const int px = bx + lx, py = by + ly;
        const bool live = has && laneOk && px < T.x1 && py < T.y1;
        const int gx = (px - B.sx0) * nx + sm, gy = (py - B.sy0) * ny + sn;
        const int bgx0 = (bx - B.sx0) * nx, bgx1 = (min(bx + bw, T.x1) - B.sx0) * nx;
        const int bgy0 = (by - B.sy0) * ny, bgy1 = (min(by + bh, T.y1) - B.sy0) * ny;
        float sx = 0.f, sy = 0.f, stime = 0.f;
        bool lodOk = false;
        if (live) {
            const uchar4 pat = pixPattern[(size_t)(py - fr->extY0) * fr->extW + (px - fr->extX0)];
            sx = (float)px + pos2d[2 * ((size_t)pat.x * ns + si)];
            sy = (float)py + pos2d[2 * ((size_t)pat.x * ns + si) + 1];
            // time = (close - open) * t + open (imagepixel.cpp:353-355)
            stime = (closetime - opentime) * pos1d[(size_t)pat.y * ns + si] + opentime;
            const float lod = pos1d[(size_t)pat.z * ns + si];
            lodOk = !(0.0f > lod || lod >= 1.0f);
        }
        float occlZ = FLT_MAX, wu = 0.f, wv = 0.f;
        float sD = FLT_MAX;               // midpoint: the occluding hit's depth
        uint32_t win = 0xffffffffu, head = 0xffffffffu;
        int nh = 0;
        const unsigned long long liveM = __ballot(live);
        smp[wave][lane] = make_float4(sx, sy, stime, lodOk ? 1.0f : 0.0f);

        for (uint32_t c0 = 0; c0 < n;) {
            const int cnt = (int)((n - c0) < (uint32_t)RH_MB_RECS ? (n - c0) : (uint32_t)RH_MB_RECS);
            __syncthreads();
            for (int q = tid; q < cnt * 8; q += NW * 64) {
                const uint32_t id = L[c0 + q / 8];
                recs[q] = mprecMB[(size_t)id * 8 + (q & 7)];
                if ((q & 7) == 0) ids[q / 8] = id;
            }
            __syncthreads();
            // per record: key bounds, sub-bound count and time step
            if (tid < cnt) {
                const float4* R = &recs[tid * 8];
                const uint32_t flags = __float_as_uint(R[6].x);
                int d = 1;
                float dtv = 0.0f;
                if (flags & RH_MB_MOVING) {
                    const float4 r0 = R[0], r1 = R[1], r2 = R[2], r3 = R[3], r4 = R[4], r5 = R[5];
                    const f3d K0[4] = {mk3(r0.x, r0.y, r0.z), mk3(r0.w, r1.x, r1.y), mk3(r1.z, r1.w, r2.x),
                                       mk3(r2.y, r2.z, r2.w)};
                    const f3d K1[4] = {mk3(r3.x, r3.y, r3.z), mk3(r3.w, r4.x, r4.y), mk3(r4.z, r4.w, r5.x),
                                       mk3(r5.y, r5.z, r5.w)};
                    float kb[2][6];
                    mb_key_bound(K0, kb[0]);
                    mb_key_bound(K1, kb[1]);
#pragma unroll
                    for (int c = 0; c < 6; ++c) {
                        kbs[tid][c] = kb[0][c];
                        kbs[tid][6 + c] = kb[1][c];
                    }
                    d = mb_divisions(kb[0], K0[0], K1[0], timeRanges);
                    dtv = (closetime - opentime) / d;
                }
                divs[tid] = d;
                dts[tid] = dtv;
            }
            __syncthreads();
            // pair offsets, and the records whose pairs fit in one pass
            if (wave == 0) {
                constexpr int PER = RH_MB_RECS / 64;
                int dk[PER], incl = 0;
#pragma unroll
                for (int j = 0; j < PER; ++j) {
                    dk[j] = PER * lane + j < cnt ? divs[PER * lane + j] : 0;
                    incl += dk[j];
                }
                for (int off = 1; off < 64; off <<= 1) {
                    const int y = __shfl_up(incl, off);
                    if (lane >= off) incl += y;
                }
                int run = incl, fit = 0;
#pragma unroll
                for (int j = 0; j < PER; ++j) run -= dk[j];
#pragma unroll
                for (int j = 0; j < PER; ++j) {
                    offs[PER * lane + j] = run;
                    run += dk[j];
                    fit += (PER * lane + j < cnt && run <= RH_MB_PAIRS) ? 1 : 0;
                }
                if (lane == 63) offs[RH_MB_RECS] = incl;
                for (int off = 32; off > 0; off >>= 1) fit += __shfl_xor(fit, off);
                if (lane == 0) nfitShared = fit;
            }
            __syncthreads();
            const int nrec = nfitShared;
            const int npairs = offs[nrec];
            // 1. entries in pair order
            int nent = 0;
            for (int pp = 0; pp < npairs; pp += NW * 64) {
                const int p = pp + tid;
                bool keep = false;
                uint32_t ea = 0, ex = 0, ey = 0;
                float emz = 0.0f;
                float4 ebd = make_float4(0.f, 0.f, 0.f, 0.f);
                float2 ett = make_float2(-INFINITY, INFINITY);
                if (p < npairs) {
                    int lo = 0, hi = nrec - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (offs[mid] <= p) lo = mid;
                        else hi = mid - 1;
                    }
                    const int r = lo, b = p - offs[r];
                    const float4* R = &recs[r * 8];
                    const float4 r6 = R[6];
                    const uint32_t flags = __float_as_uint(r6.x);
                    if (!(flags & RH_MB_MOVING)) {
                        // RenderMPG_Static's candidate samples: the bound's
                        // pixel range with the first/last pixel sub-sample clip
                        const float4 r7 = R[7];
                        const float bminx = r7.x, bmaxx = r7.y, bminy = r7.z, bmaxy = r7.w;
                        int eX = lceil_i(bmaxx), eY = lceil_i(bmaxy);
                        if (eX > B.sx1) eX = B.sx1;
                        if (eY > B.sy1) eY = B.sy1;
                        int sX = floor_i(bminx), sY = floor_i(bminy);
                        if (sY < B.sy0) sY = B.sy0;
                        if (sX < B.sx0) sX = B.sx0;
                        if (sX < eX && sY < eY) {
                            const int im = (bminx < sX) ? 0 : floor_i((bminx - sX) * nx);
                            const int in = (bminy < sY) ? 0 : floor_i((bminy - sY) * ny);
                            const int em = (bmaxx > eX) ? nx : lceil_i((bmaxx - (eX - 1)) * nx);
                            const int en = (bmaxy > eY) ? ny : lceil_i((bmaxy - (eY - 1)) * ny);
                            const int gx0 = max((sX - B.sx0) * nx + im, tgx0), gx1 = min((eX - 1 - B.sx0) * nx + em, tgx1);
                            const int gy0 = max((sY - B.sy0) * ny + in, tgy0), gy1 = min((eY - 1 - B.sy0) * ny + en, tgy1);
                            keep = gx0 < gx1 && gy0 < gy1;
                            ea = (uint32_t)r | (0xffu << 8) | ((uint32_t)ns << 24);
                            ex = (uint32_t)gx0 | ((uint32_t)gx1 << 16);
                            ey = (uint32_t)gy0 | ((uint32_t)gy1 << 16);
                            // CalculateBound's min z: min(min(min(A,B),C),D), A..D = P0,P1,P3,P2
                            const float z0 = R[0].z, z1 = R[1].y, z2 = R[2].x, z3 = R[2].w;
                            emz = fmin_ref(fmin_ref(fmin_ref(z0, z1), z3), z2);
                            ebd = r7;
                        }
                    } else {
                        const float* kb0 = &kbs[r][0];
                        const float* kb1 = &kbs[r][6];
                        const float T0 = r6.z, T1 = r6.w;
                        const int d = divs[r];
                        const float dt = dts[r];
                        float time = opentime + dt, tp = 0.0f;
                        for (int i = 0; i < b; ++i) {
                            tp = time;
                            time += dt;
                        }
                        const float tc = time;
                        const float mixc = (tc - T0) / (T1 - T0);
                        const float mixp = (tp - T0) / (T1 - T0);
                        float sb[6];
#pragma unroll
                        for (int c = 0; c < 6; ++c) {
                            const float cur = kb0[c] + mixc * (kb1[c] - kb0[c]);
                            const float prev = b == 0 ? kb0[c] : kb0[c] + mixp * (kb1[c] - kb0[c]);
                            sb[c] = c < 3 ? fmin_ref(prev, cur) : fmax_ref(prev, cur);
                        }
                        const float time0 = tc - dt;
                        const float time1 = (b != d - 1) ? (tc + dt) - dt : closetime;
                        if (!(time1 < opentime || time0 > closetime) && !(sb[2] > clipFar || sb[5] < clipNear)) {
                            int indexT0, indexT1;
                            if (fastShutter) {
                                indexT0 = 0;
                                indexT1 = ns;
                            } else {
                                indexT0 = lfloor_i((time0 - opentime) * timePerSample);
                                if (indexT0 < 0) indexT0 = 0;
                                indexT1 = lceil_i((time1 - opentime) * timePerSample);
                            }
                            int eX = lceil_i(sb[3]), eY = lceil_i(sb[4]);
                            if (eX > B.sx1) eX = B.sx1;
                            if (eY > B.sy1) eY = B.sy1;
                            int sX = floor_i(sb[0]), sY = floor_i(sb[1]);
                            if (sY < B.sy0) sY = B.sy0;
                            if (sX < B.sx0) sX = B.sx0;
                            if (sX < eX && sY < eY && indexT0 < ns) {
                                int idxEnd = indexT1 > indexT0 + 1 ? indexT1 : indexT0 + 1;
                                if (idxEnd > ns) idxEnd = ns;
                                const int gx0 = max((sX - B.sx0) * nx, tgx0), gx1 = min((eX - B.sx0) * nx, tgx1);
                                const int gy0 = max((sY - B.sy0) * ny, tgy0), gy1 = min((eY - B.sy0) * ny, tgy1);
                                keep = gx0 < gx1 && gy0 < gy1;
                                ea = (uint32_t)r | ((uint32_t)b << 8) | ((uint32_t)indexT0 << 16) |
                                     ((uint32_t)idxEnd << 24);
                                ex = (uint32_t)gx0 | ((uint32_t)gx1 << 16);
                                ey = (uint32_t)gy0 | ((uint32_t)gy1 << 16);
                                emz = sb[2];
                                ebd = make_float4(sb[0], sb[3], sb[1], sb[4]);
                                ett = make_float2(time0, time1);
                            }
                        }
                    }
                }
                const unsigned long long m = __ballot(keep);
                if (lane == 0) wtot[wave] = (uint32_t)__popcll(m);
                __syncthreads();
                int before = nent, total = 0;
                for (int w = 0; w < NW; ++w) {
                    const int v = (int)wtot[w];
                    if (w < wave) before += v;
                    total += v;
                }
                if (keep) {
                    const int pos = before + __popcll(m & lanesBelow);
                    E0[pos] = make_uint4(ea, ex, ey, __float_as_uint(emz));
                    E1[pos] = ebd;
                    E2[pos] = ett;
                }
                nent += total;
                __syncthreads();
            }
            // 2. the block's visits
            if (has) {
                int fill = 0, g = 0;
                for (;;) {
                    while (fill < 64 && g < nent) {
                        const int e = g + lane;
                        bool rel = false;
                        if (e < nent) {
                            const uint4 h = E0[e];
                            const uint32_t xr = h.y, yr = h.z;
                            rel = ((int)(xr & 0xffffu) < bgx1) & ((int)(xr >> 16) > bgx0) &
                                  ((int)(yr & 0xffffu) < bgy1) & ((int)(yr >> 16) > bgy0);
                        }
                        const unsigned long long m = __ballot(rel);
                        if (rel) buf[wave][fill + __popcll(m & lanesBelow)] = (uint16_t)e;
                        fill += __popcll(m);
                        g += 64;
                    }
                    if (fill == 0) break;
                    const int nq = fill < 64 ? fill : 64;
                    // Per entry (one lane each): the lanes it visits, its
                    // rectangle's lanes and-ed with its index window's (every
                    // visit is a sample test), and of those the ones passing
                    // the tests that do not depend on earlier visits: the
                    // sub-bound's time window, its bound (Contains2D) and LOD
                    // (bucketprocessor.cpp:1393-1420; a static entry's times
                    // are unbounded).  Then the 64 x 64 bit matrix is
                    // transposed, so each lane holds the entries that pass
                    // for its sample.
                    unsigned long long pass = 0;
                    if (lane < nq) {
                        const int e = buf[wave][lane];
                        const uint4 h = E0[e];
                        const float4 bd = E1[e];
                        const float2 tt = E2[e];
                        const unsigned long long vis = block_rect_lanes(make_uint2(h.y, h.z), bgx0, bgy0, BW, BH, rowRep) &
                                                       winLanes[(h.x >> 24)] & ~winLanes[(h.x >> 16) & 0xffu] & liveM;
                        nTest += (uint32_t)__popcll(vis);
                        for (unsigned long long v = vis; v != 0ull; v &= v - 1ull) {
                            const int j = __ffsll((long long)v) - 1;
                            const float4 sp = smp[wave][j];    // x, y, time, LOD test
                            const bool ok = !(sp.z < tt.x || sp.z > tt.y) &
                                            !(sp.x < bd.x || sp.x > bd.y || sp.y < bd.z || sp.y > bd.w) & (sp.w != 0.0f);
                            pass |= (unsigned long long)ok << j;
                        }
                    }
                    const unsigned long long vm = transpose64(pass, lane);
                    // Two passes over the visits of these entries.  A: every
                    // visit's tests but the occlusion test, 64 visits a step
                    // whichever samples they belong to, taken in rank-major
                    // order (each lane's first visit, then each lane's
                    // second, ...), results into a ring in LDS.  B: once a
                    // rank is complete, each lane applies its visit of that
                    // rank in order: occlusion test, counters, StoreSample.
                    const int cnt = __popcll(vm);
                    int maxc = cnt;
                    for (int off = 32; off > 0; off >>= 1) maxc = max(maxc, __shfl_xor(maxc, off));
                    int total = 0;
                    for (int r = 0; r < maxc; ++r) {
                        const unsigned long long lm = __ballot(cnt > r);
                        if (lane == 0) {
                            rankStart[wave][r] = total;
                            rankLanes[wave][r] = lm;
                        }
                        total += __popcll(lm);
                    }
                    if (lane == 0) rankStart[wave][maxc] = total;
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    const uint32_t vmLo = (uint32_t)vm, vmHi = (uint32_t)(vm >> 32);
                    int done = 0, rB = 0;
                    while (done < total) {
                        // A: visits [done, done + 64)
                        {
                            const int p = done + lane;
                            const bool pv = p < total;
                            int rk = 0;
                            {
                                int lo = 0, hi = maxc - 1;
                                while (lo < hi) {
                                    const int mid = (lo + hi + 1) >> 1;
                                    if (rankStart[wave][mid] <= p) lo = mid;
                                    else hi = mid - 1;
                                }
                                rk = lo;
                            }
                            const int own = pv ? select_bit(rankLanes[wave][rk], p - rankStart[wave][rk]) : lane;
                            const unsigned long long vmo =
                                ((unsigned long long)(uint32_t)__shfl((int)vmHi, own) << 32) | (uint32_t)__shfl((int)vmLo, own);
                            const float osx = __shfl(sx, own), osy = __shfl(sy, own), ost = __shfl(stime, own);
                            uint32_t fl = 0;
                            float D = 0.f, u = 0.f, v = 0.f;
                            int q = 0;
                            if (pv) {
                                q = select_bit(vmo, rk);
                                const int e = buf[wave][q];
                                const uint32_t a = E0[e].x;
                                const int r = (int)(a & 0xffu), b = (int)((a >> 8) & 0xffu);
                                const float4* R = &recs[r * 8];
                                const float4 r6 = R[6];
                                const uint32_t flags = __float_as_uint(r6.x);
                                const uint32_t code = flags & 0xffu;
                                const bool degen = (flags >> 8) & 1u;
                                const float4 r0 = R[0], r1 = R[1], r2 = R[2];
                                const f3d K0[4] = {mk3(r0.x, r0.y, r0.z), mk3(r0.w, r1.x, r1.y),
                                                   mk3(r1.z, r1.w, r2.x), mk3(r2.y, r2.z, r2.w)};
                                bool tight = true;
                                f3d pts[4] = {K0[0], K0[1], K0[2], K0[3]};
                                if (b != 0xff) {
                                    // CqMicroPolygonMotion::Sample: the key span of
                                    // the time, the tight bound and the vertices there
                                    const float* kb0 = &kbs[r][0];
                                    const float* kb1 = &kbs[r][6];
                                    const float T0 = r6.z, T1 = r6.w;
                                    int iIndex = 0;
                                    float Fraction = 0.0f;
                                    bool Exact = true;
                                    if (ost > T0) {
                                        if (ost >= T1) {
                                            iIndex = 1;
                                        } else {
                                            Fraction = (ost - T0) / (T1 - T0);
                                            Exact = (T0 == ost);
                                        }
                                    }
                                    const float* kbi = iIndex ? kb1 : kb0;
                                    float tb[4];
                                    if (Exact) {
                                        tb[0] = kbi[0]; tb[1] = kbi[3]; tb[2] = kbi[1]; tb[3] = kbi[4];
                                    } else {
                                        const float F1 = 1 - Fraction;
                                        tb[0] = F1 * kb0[0] + Fraction * kb1[0];
                                        tb[1] = F1 * kb0[3] + Fraction * kb1[3];
                                        tb[2] = F1 * kb0[1] + Fraction * kb1[1];
                                        tb[3] = F1 * kb0[4] + Fraction * kb1[4];
                                    }
                                    tight = !(osx < tb[0] || osx > tb[1] || osy < tb[2] || osy > tb[3]);
                                    const float4 r3 = R[3], r4 = R[4], r5 = R[5];
                                    const f3d K1[4] = {mk3(r3.x, r3.y, r3.z), mk3(r3.w, r4.x, r4.y),
                                                       mk3(r4.z, r4.w, r5.x), mk3(r5.y, r5.z, r5.w)};
                                    // a key's vertices when the time is exact, else
                                    // (1 - F) K0 + F K1; selected per component
                                    const float F1 = 1.0f - Fraction;
                                    const bool k1 = iIndex != 0;
#pragma unroll
                                    for (int k = 0; k < 4; ++k) {
                                        const f3d mixed = mk3((F1 * K0[k].x) + (Fraction * K1[k].x),
                                                              (F1 * K0[k].y) + (Fraction * K1[k].y),
                                                              (F1 * K0[k].z) + (Fraction * K1[k].z));
                                        pts[k].x = Exact ? (k1 ? K1[k].x : K0[k].x) : mixed.x;
                                        pts[k].y = Exact ? (k1 ? K1[k].y : K0[k].y) : mixed.y;
                                        pts[k].z = Exact ? (k1 ? K1[k].z : K0[k].z) : mixed.z;
                                    }
                                }
                                if (tight) {
                                    const MBHit h = quad_hit(pts[0], pts[1], pts[2], pts[3], code, degen, osx, osy);
                                    if (h.in) {
                                        fl = 2u | (h.linear ? 4u : 0u);
                                        D = h.D; u = h.u; v = h.v;
                                    }
                                }
                                const int slot = p & (RH_MB_RING - 1);
                                ringCode[wave][slot] = (uint32_t)q | (fl << 8);
                                ringD[wave][slot] = D;
                                ringU[wave][slot] = u;
                                ringV[wave][slot] = v;
                            }
                        }
                        done = min(done + 64, total);
                        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        // B: the ranks now complete, in order
                        while (rB < maxc && rankStart[wave][rB + 1] <= done) {
                            const bool act = cnt > rB;
                            const unsigned long long lm = __ballot(act);
                            bool hit = false, opaque = false;
                            float D = 0.f, u = 0.f, v = 0.f;
                            uint32_t hid = 0;
                            int hgrid = 0;
                            if (act) {
                                const int slot = (rankStart[wave][rB] + __popcll(lm & lanesBelow)) & (RH_MB_RING - 1);
                                const uint32_t rc = ringCode[wave][slot];
                                const int e = buf[wave][rc & 0xffu];
                                const uint4 h = E0[e];
                                const float mz = __uint_as_float(h.w);
                                const int r = (int)(h.x & 0xffu);
                                const float4 r6 = recs[r * 8 + 6];
                                opaque = (__float_as_uint(r6.x) >> 10) & 1u;
                                hid = ids[r];
                                hgrid = __float_as_int(r6.y);
                                if (!(cullH && mz > occlZ)) {
                                    ++nBound;
                                    if (rc & 0x200u) {
                                        ++nIn;
                                        if (!(rc & 0x400u)) ++nIn2;
                                        D = ringD[wave][slot];
                                        u = ringU[wave][slot];
                                        v = ringV[wave][slot];
                                        hit = !cullH || !(occlZ <= D);
                                    }
                                }
                            }
                            // StoreSample (bucketprocessor.cpp:1471-1569); the
                            // midpoint case keeps the nearest opaque hit and
                            // occludes with the second nearest (:1502-1525)
                            nHit += hit ? 1u : 0u;
                            const bool opq = hit && opaque && cullH;
                            bool store = opq;
                            if (midS) {
                                const float prevZ = win != 0xffffffffu ? sD : FLT_MAX;
                                const bool keep = !(prevZ < D);
                                store = opq && keep;
                                occlZ = opq ? (keep ? prevZ : D) : occlZ;
                                sD = store ? D : sD;
                            } else {
                                occlZ = opq ? D : occlZ;
                            }
                            win = store ? hid : win;
                            wu = store ? u : wu;
                            wv = store ? v : wv;
                            const bool app = hit && !opq && !depthOnly;
                            const unsigned long long am = __ballot(app);
                            if (am) {
                                const int na = __popcll(am);
                                if (na > wbLeft) {
                                    uint32_t base = 0;
                                    if (lane == 0) base = atomicAdd(hitCounter, 64u);
                                    wbBase = __shfl(base, 0);
                                    wbLeft = 64;
                                }
                                if (app) {
                                    const uint32_t slot = wbBase + (uint32_t)(64 - wbLeft) + (uint32_t)__popcll(am & lanesBelow);
                                    if (slot < fr->hitCapacity) {
                                        hitData[slot] = make_float4(u, v, D, __uint_as_float(hid));
                                        hitNext[2 * (size_t)slot] = head;
                                        hitNext[2 * (size_t)slot + 1] = grids[hgrid].prim;
                                        head = slot;
                                        ++nh;
                                    }
                                }
                                wbLeft -= na;
                            }
                            ++rB;
                        }
                    }
                    // keep what the buffer holds past the first 64
                    const int rest = fill - nq;
                    const uint16_t keepE = lane < rest ? buf[wave][64 + lane] : (uint16_t)0;
                    if (lane < rest) buf[wave][lane] = keepE;
                    fill = rest;
                }
            }
            c0 += (uint32_t)nrec;
        }
        if (live) {
            const size_t sidx = ((size_t)(py - fr->extY0) * fr->extW + (px - fr->extX0)) * ns + si;
            float4 a = make_float4(0.f, 0.f, 0.f, FLT_MAX), b = make_float4(0.f, 0.f, 0.f, __uint_as_float(win));
            if (win != 0xffffffffu) {
                const float4 r6w = mprecMB[(size_t)win * 8 + 6];
                const PrimDev& P = prims[grids[__float_as_int(r6w.y)].prim];
                shade_hit(P, wu, wv, midS ? sD : occlZ, win, a, b);
                finish_occluding(fr, P, nh > 0, depthOnly, occlZ, a, b);
            }
            samples[2 * sidx] = a;
            samples[2 * sidx + 1] = b;
            if (frontDepth) frontDepth[sidx] = nh > 0 ? occlZ : a.w;
        }
        const bool multi = live && nh > 0;
        const unsigned long long cm = __ballot(multi);
        if (cm) {
            uint32_t base = 0;
            if (lane == __ffsll((long long)cm) - 1) base = atomicAdd(&hitCounter[2], (uint32_t)__popcll(cm));
            base = __shfl(base, __ffsll((long long)cm) - 1);
            if (multi) {
                const size_t sidx = ((size_t)(py - fr->extY0) * fr->extW + (px - fr->extX0)) * ns + si;
                const uint32_t e = base + (uint32_t)__popcll(cm & lanesBelow);
                combineList[e] = make_uint4((uint32_t)sidx, (uint32_t)(sidx >> 32), head, (uint32_t)nh);
            }
        }
    }
    }   // task < ntasks
    for (int off = 32; off > 0; off >>= 1) {
        nTest += __shfl_xor(nTest, off);
        nBound += __shfl_xor(nBound, off);
        nHit += __shfl_xor(nHit, off);
        nIn += __shfl_xor(nIn, off);
        nIn2 += __shfl_xor(nIn2, off);
    }
    if (lane == 0) {
        wsum[wave][0] = nTest; wsum[wave][1] = nBound; wsum[wave][2] = nHit; wsum[wave][3] = nIn;
        wsum[wave][4] = nIn2;
    }
    __syncthreads();
    if (tid < 5) {
        uint32_t v = 0;
        for (int w = 0; w < NW; ++w) v += wsum[w][tid];
        rowCounts[(size_t)(rowBase + blockIdx.x) * 5 + tid] = v;
    }
}

// ---------------------------------------------------------------------------
// Sampling with depth of field
// ---------------------------------------------------------------------------
// GetCircleOfConfusion (renderer.h:401-406): raster radii at a camera depth.
__device__ __forceinline__ float2 dof_coc(const FrameDev* __restrict__ fr, float depth) {
    const float c = fr->dofMult * fabsf(1.0f / depth - fr->dofOneOverF);
    return make_float2(fr->dofScaleX * c, fr->dofScaleY * c);
}

// CqMicroPolygon::Sample with depth of field (micropolygon.cpp:1561-1593)
// after dofSampleInBound: the four grid-order vertices moved against the
// sample's lens offset, cachePointInPolyTest's edges and CqInvBilinear's
// set-up on them (micropolygon.cpp:1346-1392, bilinear.h:260-275; the
// linear test of the moved quad decides one or two Newton steps), fContains
// (:1295-1342), the inverse bilinear (bilinear.h:277-305) and the depth
// bilerp (:228-236).
__device__ __forceinline__ bool dof_quad_hit(const float* qx, const float* qy, const float4 z, uint32_t code,
                                             bool degen, float px, float py, float& u, float& v, float& D,
                                             bool& linear) {
    float ox[4], oy[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int o = i < 3 ? (code >> (2 * (i + 1))) & 3 : code & 3;
        ox[i] = pick4(o, qx[0], qx[1], qx[2], qx[3]);
        oy[i] = pick4(o, qy[0], qy[1], qy[2], qy[3]);
    }
    float YM[4], XM[4], X[4], Y[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int j = (i + 3) & 3;
        YM[i] = ox[i] - ox[j];
        XM[i] = oy[i] - oy[j];
        X[i] = ox[j];
        Y[i] = oy[j];
    }
    if (degen) {
#pragma unroll
        for (int i = 2; i < 4; ++i) {
            YM[i] = ox[3] - ox[1];
            XM[i] = oy[3] - oy[1];
            X[i] = ox[1];
            Y[i] = oy[1];
        }
    }
    const float t0 = ((py - Y[0]) * YM[0]) - ((px - X[0]) * XM[0]);
    const float t1 = ((py - Y[1]) * YM[1]) - ((px - X[1]) * XM[1]);
    const float t2 = ((py - Y[2]) * YM[2]) - ((px - X[2]) * XM[2]);
    const float t3 = ((py - Y[3]) * YM[3]) - ((px - X[3]) * XM[3]);
    if (t0 <= 0 || t1 <= 0 || t2 < 0 || t3 < 0) return false;
    const float Ax = qx[0], Ay = qy[0];
    const float Ex = qx[1] - qx[0], Ey = qy[1] - qy[0];
    const float Fx = qx[2] - qx[0], Fy = qy[2] - qy[0];
    const float Gx = ((-Ex) - qx[2]) + qx[3], Gy = ((-Ey) - qy[2]) + qy[3];
    const float patchSize = fmax_ref(fmax_ref(fabsf(Fx), fabsf(Fy)), fmax_ref(fabsf(Ex), fabsf(Ey)));
    const float irregularity = fmax_ref(fabsf(Gx), fabsf(Gy));
    linear = (double)irregularity < 1e-2 * (double)patchSize;
    u = 0.5f;
    v = 0.5f;
    {
        const float M1x = Ex + Gx * v, M1y = Ey + Gy * v;
        const float M2x = Fx + Gx * u, M2y = Fy + Gy * u;
        const float bx = (((Ax + Ex * u) + Fx * v) + (Gx * u) * v) - px;
        const float by = (((Ay + Ey * u) + Fy * v) + (Gy * u) * v) - py;
        float det = M1x * M2y - M1y * M2x;
        det = 1 / det;
        u = u - det * (bx * M2y - by * M2x);
        v = v - det * (-(bx * M1y - by * M1x));
    }
    if (!linear) {
        const float M1x = Ex + Gx * v, M1y = Ey + Gy * v;
        const float M2x = Fx + Gx * u, M2y = Fy + Gy * u;
        const float bx = (((Ax + Ex * u) + Fx * v) + (Gx * u) * v) - px;
        const float by = (((Ay + Ey * u) + Fy * v) + (Gy * u) * v) - py;
        float det = M1x * M2y - M1y * M2x;
        if (det != 0) det = 1 / det;
        u = u - det * (bx * M2y - by * M2x);
        v = v - det * (-(bx * M1y - by * M1x));
    }
    const float w0 = (1 - v) * (1 - u), w1 = (1 - v) * u, w2 = v * (1 - u), w3 = v * u;
    D = ((w0 * z.x + w1 * z.y) + w2 * z.z) + w3 * z.w;
    return true;
}

#define RH_DOF_CH 128
#define RH_DOF_WAVES 4
// RenderMPG_MBOrDof for static micropolygons with depth of field
// (bucketprocessor.cpp:1221-1469, the UsingDof branches).  A wave block is
// one lens cell of an 8 x 8 pixel tile of a task: lane = pixel, and the
// pixel's sample is the one visited through that cell (GetDofOffsetIndex:
// the pixel's shuffle table entry of the cell), whose lens offset is the
// cell's entry of the pixel's dof pattern.  Per block and micropolygon, in
// list order: the bound shifted by the cell's bound of offsets times the
// larger circle of confusion at the micropolygon's depths gives the pixel
// window (every pixel of it counts into SPL_count); the shifted bound's
// Contains2D and the LOD range are order-free; then each lane, in list
// order: occlusion skip, SPL_bound_hits, dofSampleInBound (micropolygon.cpp:
// 1531-1550) against the micropolygon's bound, the moved-quad hit test and
// StoreSample.  One workgroup of four waves per task.  STORE: StoreSample's
// mode (RH_STORE_*).
template <int STORE>
__global__ void __launch_bounds__(RH_DOF_WAVES * 64)
rh_sample_dof(const FrameDev* __restrict__ fr, const BucketDev* __restrict__ buckets, const TaskDev* __restrict__ tasks,
              int itemBase, const unsigned long long* __restrict__ capStart, const uint32_t* __restrict__ listCount,
              const uint32_t* __restrict__ lists, const float4* __restrict__ mprec4, const GridDev* __restrict__ grids,
              const PrimDev* __restrict__ prims, const float* __restrict__ pos2d, const float* __restrict__ pos1d,
              const uchar4* __restrict__ pixPattern, const uint8_t* __restrict__ shuffled,
              const float2* __restrict__ lensTab, const uint8_t* __restrict__ pixDof, const float4* __restrict__ cells,
              float4* __restrict__ samples, float4* __restrict__ hitData, uint32_t* __restrict__ hitNext,
              uint32_t* __restrict__ hitCounter, uint4* __restrict__ combineList, float* __restrict__ frontDepth,
              uint32_t* __restrict__ rowCounts, int rowBase, uint32_t idCut, const uint32_t* __restrict__ itemCut,
              const uint32_t* __restrict__ taskList, int depthOnly) {
    __shared__ float4 recs[RH_DOF_CH * 9];
    __shared__ uint8_t matched[RH_DOF_WAVES][RH_DOF_CH];
    __shared__ unsigned long long matM[RH_DOF_WAVES][RH_DOF_CH];
    __shared__ float4 matB[RH_DOF_WAVES][RH_DOF_CH];
    __shared__ uint32_t wsum[RH_DOF_WAVES][5];
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ns = fr->ns;
    const int task = taskList ? (int)taskList[itemBase + blockIdx.x] : itemBase + (int)blockIdx.x;
    const TaskDev T = tasks[task];
    const BucketDev B = buckets[T.bucket];
    constexpr bool cullH = STORE != RH_STORE_LIST_ALL, midS = STORE == RH_STORE_MIDPOINT;
    const int tilesX = (T.x1 - T.x0 + 7) >> 3, tilesY = (T.y1 - T.y0 + 7) >> 3;
    const int nblk = tilesX * tilesY * ns;
    const uint32_t* L = lists + capStart[task];
    const uint32_t n = list_cut(L, listCount[task], itemCut ? itemCut[itemBase + blockIdx.x] : idCut);
    const unsigned long long rowRep = 0x0101010101010101ull;   // bit 8r of every row r
    const int lx = lane & 7, ly = lane >> 3;
    uint32_t nTest = 0, nBound = 0, nHit = 0, nIn = 0, nIn2 = 0;
    uint32_t wbBase = 0;
    int wbLeft = 0;

    // every wave runs the same number of passes (the staging barriers)
    for (int p0 = 0; p0 < nblk; p0 += RH_DOF_WAVES) {
        const int blk = p0 + wave;
        const bool has = blk < nblk;                     // uniform across the wave
        const int tile = has ? blk / ns : 0, cell = has ? blk - tile * ns : 0;
        const int tx0 = T.x0 + (tile % tilesX) * 8, ty0 = T.y0 + (tile / tilesX) * 8;
        const int px = tx0 + lx, py = ty0 + ly;
        const bool live = has && px < T.x1 && py < T.y1;
        float sx = 0.f, sy = 0.f, lox = 0.f, loy = 0.f;
        bool lodOk = false;
        size_t sidx = 0;
        if (live) {
            const size_t pix = (size_t)(py - fr->extY0) * fr->extW + (px - fr->extX0);
            const uchar4 pat = pixPattern[pix];
            const int si = shuffled[(size_t)pat.w * ns + cell];
            sx = (float)px + pos2d[2 * ((size_t)pat.x * ns + si)];
            sy = (float)py + pos2d[2 * ((size_t)pat.x * ns + si) + 1];
            const float lod = pos1d[(size_t)pat.z * ns + si];
            lodOk = !(0.0f > lod || lod >= 1.0f);
            const float2 lo = lensTab[(size_t)pixDof[pix] * ns + cell];
            lox = lo.x;
            loy = lo.y;
            sidx = pix * ns + si;
        }
        const float4 DB = has ? cells[cell] : make_float4(0.f, 0.f, 0.f, 0.f);
        const unsigned long long liveM = __ballot(live), inM = __ballot(live && lodOk);
        const int BX0 = tx0 - B.sx0, BY0 = ty0 - B.sy0;
        float occlZ = FLT_MAX, wu = 0.f, wv = 0.f;
        float sD = FLT_MAX;               // midpoint: the occluding hit's depth
        uint32_t win = 0xffffffffu, head = 0xffffffffu;
        int nh = 0;
        for (uint32_t c0 = 0; c0 < n; c0 += RH_DOF_CH) {
            const int cnt = (int)((n - c0) < (uint32_t)RH_DOF_CH ? (n - c0) : (uint32_t)RH_DOF_CH);
            __syncthreads();
            if (tid < cnt) {
                const uint32_t id = L[c0 + tid];
                const float4* src = mprec4 + (size_t)id * (RH_MPREC_FLOATS / 4);
                const float4 a = src[0], b = src[1], zz = src[2], bd = src[3], r4 = src[4];
                // CacheHitTestValues with DoF (micropolygon.cpp:1394-1421):
                // each vertex's CoC, their component-wise least and largest
                const float2 c0v = dof_coc(fr, zz.x), c1v = dof_coc(fr, zz.y), c2v = dof_coc(fr, zz.z),
                             c3v = dof_coc(fr, zz.w);
                const float cminx = fmin_ref(fmin_ref(c0v.x, c1v.x), fmin_ref(c2v.x, c3v.x));
                const float cminy = fmin_ref(fmin_ref(c0v.y, c1v.y), fmin_ref(c2v.y, c3v.y));
                const float cmaxx = fmax_ref(fmax_ref(c0v.x, c1v.x), fmax_ref(c2v.x, c3v.x));
                const float cmaxy = fmax_ref(fmax_ref(c0v.y, c1v.y), fmax_ref(c2v.y, c3v.y));
                // the larger CoC at the bound's depths (CalculateBound's
                // order: A, B, C = vertex 3, D = vertex 2)
                const float bmaxz = fmax_ref(fmax_ref(fmax_ref(zz.x, zz.y), zz.w), zz.z);
                const float2 cz0 = dof_coc(fr, r4.x), cz1 = dof_coc(fr, bmaxz);
                float4* R = &recs[tid * 9];
                R[0] = a; R[1] = b; R[2] = zz; R[3] = bd;
                R[4] = make_float4(r4.x, r4.y, r4.z, __uint_as_float(id));
                R[5] = make_float4(c0v.x, c0v.y, c1v.x, c1v.y);
                R[6] = make_float4(c2v.x, c2v.y, c3v.x, c3v.y);
                R[7] = make_float4(cminx, cminy, cmaxx, cmaxy);
                R[8] = make_float4(fmax_ref(cz0.x, cz1.x), fmax_ref(cz0.y, cz1.y),
                                   (bmaxz < fr->clipNear || r4.x > fr->clipFar) ? 1.0f : 0.0f, 0.0f);
            }
            __syncthreads();
            if (!has) continue;
            float zmax = live ? occlZ : -FLT_MAX;
            for (int off = 32; off > 0; off >>= 1) zmax = fmaxf(zmax, __shfl_xor(zmax, off));
            // (1) the chunk's micropolygons whose window for this cell meets
            // the tile, in list order, with their lanes and shifted bounds;
            // those beyond every lane's opaque depth are only counted
            int nm = 0;
            for (int j0 = 0; j0 < cnt; j0 += 64) {
                const int j = j0 + lane;
                unsigned long long m = 0ull;
                float4 sb = make_float4(0.f, 0.f, 0.f, 0.f);
                bool ok = false;
                if (j < cnt) {
                    const float4 bd = recs[j * 9 + 3], r8 = recs[j * 9 + 8];
                    if (r8.z == 0.0f) {            // outside the clipping range: skipped
                        const float bminx = bd.x - DB.z * r8.x, bmaxx = bd.y - DB.x * r8.x;
                        const float bminy = bd.z - DB.w * r8.y, bmaxy = bd.w - DB.y * r8.y;
                        int eX = lceil_i(bmaxx), eY = lceil_i(bmaxy);
                        if (eX > B.sx1) eX = B.sx1;
                        if (eY > B.sy1) eY = B.sy1;
                        int sX = floor_i(bminx), sY = floor_i(bminy);
                        if (sY < B.sy0) sY = B.sy0;
                        if (sX < B.sx0) sX = B.sx0;
                        if (sX < eX && sY < eY) {
                            const uint2 mt = make_uint2((uint32_t)(sX - B.sx0) | ((uint32_t)(eX - B.sx0) << 16),
                                                        (uint32_t)(sY - B.sy0) | ((uint32_t)(eY - B.sy0) << 16));
                            m = block_rect_lanes(mt, BX0, BY0, 8, 8, rowRep) & liveM;
                        }
                        sb = make_float4(bminx, bmaxx, bminy, bmaxy);
                    }
                    nTest += (uint32_t)__popcll(m);
                    ok = m != 0ull && !(cullH && recs[j * 9 + 4].x > zmax);
                }
                const unsigned long long msk = __ballot(ok);
                if (ok) {
                    const int slot = nm + __popcll(msk & ((1ull << lane) - 1ull));
                    matched[wave][slot] = (uint8_t)j;
                    matM[wave][slot] = m & inM;
                    matB[wave][slot] = sb;
                }
                nm += __popcll(msk);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (int i0 = 0; i0 < nm; i0 += 64) {
                const int ni = (nm - i0) < 64 ? (nm - i0) : 64;
                // (2) order-free: the lane's sample in the micropolygon's
                // window, inside its shifted bound, LOD in range
                unsigned long long pb = 0ull;
                for (int i = 0; i < ni; ++i) {
                    const unsigned long long mm = matM[wave][i0 + i];
                    const float4 sb = matB[wave][i0 + i];
                    const bool inB = !(sx < sb.x) & !(sx > sb.y) & !(sy < sb.z) & !(sy > sb.w);
                    const bool in = ((mm >> lane) & 1ull) != 0ull && inB;
                    pb |= (unsigned long long)in << i;
                }
                // (3) ordered: each lane its own candidates in list order
                while (__ballot(pb != 0ull)) {
                    bool hit = false;
                    float D = 0.f, u = 0.5f, v = 0.5f;
                    uint32_t hflags = 0u, hid = 0u, hgrid = 0u;
                    if (pb != 0ull) {
                        const int i = __ffsll((long long)pb) - 1;
                        pb &= pb - 1ull;
                        const float4* R = &recs[matched[wave][i0 + i] * 9];
                        const float4 r4 = R[4];
                        // occlusion skip against the current opaque hit (bucketprocessor.cpp:1396-1397)
                        if (!(cullH && r4.x > occlZ)) {
                            ++nBound;
                            // dofSampleInBound (micropolygon.cpp:1531-1550)
                            const float4 cm = R[7], bd = R[3];
                            float cMinX = sx + cm.x * lox, cMinY = sy + cm.y * loy;
                            float cMaxX = sx + cm.z * lox, cMaxY = sy + cm.w * loy;
                            if (lox < 0) { const float t = cMinX; cMinX = cMaxX; cMaxX = t; }
                            if (loy < 0) { const float t = cMinY; cMinY = cMaxY; cMaxY = t; }
                            if (!(cMinX > bd.y || cMinY > bd.w || cMaxX < bd.x || cMaxY < bd.z)) {
                                const float4 a = R[0], b = R[1], c01 = R[5], c23 = R[6];
                                // the vertices moved against the lens offset
                                const float qx[4] = {a.x - c01.x * lox, a.z - c01.z * lox, b.x - c23.x * lox,
                                                     b.z - c23.z * lox};
                                const float qy[4] = {a.y - c01.y * loy, a.w - c01.w * loy, b.y - c23.y * loy,
                                                     b.w - c23.w * loy};
                                hflags = __float_as_uint(r4.y);
                                bool lin = false;
                                if (dof_quad_hit(qx, qy, R[2], hflags & 0xffu, ((hflags >> 8) & 1u) != 0u, sx, sy,
                                                 u, v, D, lin)) {
                                    ++nIn;
                                    if (!lin) ++nIn2;
                                    // StoreSample (bucketprocessor.cpp:1471-1569): reject when occlZ <= D
                                    hit = !cullH || !(occlZ <= D);
                                    hid = __float_as_uint(r4.w);
                                    hgrid = __float_as_uint(r4.z);
                                }
                            }
                        }
                    }
                    nHit += hit ? 1u : 0u;
                    // the occluding hit; midpoint keeps the nearest opaque hit
                    // and occludes with the second nearest (:1502-1525)
                    const bool opq = hit && (hflags & (1u << 10)) != 0 && cullH;
                    const float prevZ = win != 0xffffffffu ? sD : FLT_MAX;
                    const bool keep = !midS || !(prevZ < D);
                    const bool store = opq && keep;
                    occlZ = opq ? (midS && keep ? prevZ : D) : occlZ;
                    win = store ? hid : win;
                    wu = store ? u : wu;
                    wv = store ? v : wv;
                    sD = store ? D : sD;
                    const bool app = hit && !opq && !depthOnly;
                    const unsigned long long am = __ballot(app);
                    if (am) {
                        const int na = __popcll(am);
                        if (na > wbLeft) {
                            uint32_t base = 0;
                            if (lane == 0) base = atomicAdd(hitCounter, 64u);
                            wbBase = __shfl(base, 0);
                            wbLeft = 64;
                        }
                        if (app) {
                            const uint32_t slot =
                                wbBase + (uint32_t)(64 - wbLeft) + (uint32_t)__popcll(am & ((1ull << lane) - 1ull));
                            if (slot < fr->hitCapacity) {
                                hitData[slot] = make_float4(u, v, D, __uint_as_float(hid));
                                hitNext[2 * (size_t)slot] = head;
                                hitNext[2 * (size_t)slot + 1] = grids[hgrid].prim;
                                head = slot;
                                ++nh;
                            }
                        }
                        wbLeft -= na;
                    }
                }
            }
        }
        if (!live) continue;
        float4 A = make_float4(0.f, 0.f, 0.f, FLT_MAX), Bv = make_float4(0.f, 0.f, 0.f, __uint_as_float(win));
        if (win != 0xffffffffu) {
            const float4 r4 = mprec4[(size_t)win * (RH_MPREC_FLOATS / 4) + 4];
            const PrimDev& P = prims[grids[__float_as_int(r4.z)].prim];
            shade_hit(P, wu, wv, sD, win, A, Bv);
            finish_occluding(fr, P, nh > 0, depthOnly, occlZ, A, Bv);
        }
        samples[2 * sidx] = A;
        samples[2 * sidx + 1] = Bv;
        if (frontDepth) frontDepth[sidx] = nh > 0 ? occlZ : A.w;
        const bool multi = nh > 0;
        const unsigned long long cmk = __ballot(multi);
        if (cmk) {
            uint32_t base = 0;
            if (lane == __ffsll((long long)cmk) - 1) base = atomicAdd(&hitCounter[2], (uint32_t)__popcll(cmk));
            base = __shfl(base, __ffsll((long long)cmk) - 1);
            if (multi) {
                const uint32_t e = base + (uint32_t)__popcll(cmk & ((1ull << lane) - 1ull));
                combineList[e] = make_uint4((uint32_t)sidx, (uint32_t)(sidx >> 32), head, (uint32_t)nh);
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        nTest += __shfl_xor(nTest, off);
        nBound += __shfl_xor(nBound, off);
        nHit += __shfl_xor(nHit, off);
        nIn += __shfl_xor(nIn, off);
        nIn2 += __shfl_xor(nIn2, off);
    }
    if (lane == 0) {
        wsum[wave][0] = nTest; wsum[wave][1] = nBound; wsum[wave][2] = nHit; wsum[wave][3] = nIn;
        wsum[wave][4] = nIn2;
    }
    __syncthreads();
    if (tid < 5) {
        uint32_t v = 0;
        for (int w = 0; w < RH_DOF_WAVES; ++w) v += wsum[w][tid];
        rowCounts[(size_t)(rowBase + blockIdx.x) * 5 + tid] = v;
    }
}

// ---------------------------------------------------------------------------
// Batched form: work items ordered by cost
// ---------------------------------------------------------------------------
// One workgroup.  A task's cost is (list length + 16) x blocks; a task is cut
// into k items of near-equal block ranges, at least enough that no item
// needs more than one pass (itemBlocks blocks), and more while an item would
// cost over splitCost (down to one block per item).  Items are written in
// descending order of log2(cost), so the persistent sampler takes the
// largest first; the order inside a class is unspecified (results do not
// depend on it).  Also resets the work queue.
__global__ void __launch_bounds__(1024) rh_plan_items(const FrameDev* __restrict__ fr,
                                                      const TaskDev* __restrict__ tasks, int ntasks,
                                                      const uint32_t* __restrict__ listCount, uint32_t splitCost,
                                                      int itemBlocks, uint2* __restrict__ items,
                                                      uint32_t* __restrict__ itemCount,
                                                      uint32_t* __restrict__ queueHead) {
    __shared__ uint32_t hist[33], offs[33];
    const int tid = threadIdx.x;
    const int bw = fr->blockW, bh = fr->blockH;
    if (tid < 33) hist[tid] = 0;
    __syncthreads();
    for (int pass = 0; pass < 2; ++pass) {
        for (int t = tid; t < ntasks; t += 1024) {
            const TaskDev T = tasks[t];
            const int nblk = ((T.x1 - T.x0 + bw - 1) / bw) * ((T.y1 - T.y0 + bh - 1) / bh);
            const unsigned long long cost = (unsigned long long)(listCount[t] + 16u) * (unsigned)nblk;
            int k = (nblk + itemBlocks - 1) / itemBlocks;
            while (k < nblk && cost / (unsigned long long)k > splitCost) k = (2 * k < nblk) ? 2 * k : nblk;
            // items of `per` blocks; the last may be shorter, so count them
            const int per = (nblk + k - 1) / k;
            const int nitems = (nblk + per - 1) / per;
            unsigned long long c = cost / (unsigned long long)nitems;
            if (c > 0xffffffffull) c = 0xffffffffull;
            const int bin = 32 - __clz((uint32_t)c | 1u);   // 1..32
            if (pass == 0) {
                atomicAdd(&hist[bin], (uint32_t)nitems);
            } else {
                for (int b0 = 0; b0 < nblk; b0 += per) {
                    const int nb = (nblk - b0) < per ? (nblk - b0) : per;
                    const uint32_t slot = atomicAdd(&offs[bin], 1u);
                    items[slot] = make_uint2((uint32_t)t, (uint32_t)b0 | ((uint32_t)nb << 16));
                }
            }
        }
        __syncthreads();
        if (pass == 0 && tid == 0) {
            uint32_t run = 0;
            for (int b = 32; b >= 0; --b) {
                offs[b] = run;
                run += hist[b];
            }
            itemCount[0] = run;
            queueHead[0] = 0;
        }
        __syncthreads();
    }
}

// Clears the frame's counters (graph-safe: a kernel, not a memset).
__global__ void rh_frame_begin(unsigned long long* __restrict__ counters, uint32_t* __restrict__ hitCtl) {
    const int t = threadIdx.x;
    if (t < RH_NCOUNTERS) counters[t] = 0ull;
    if (t < RH_HITCTL_WORDS) hitCtl[t] = 0u;
}

// ---------------------------------------------------------------------------
// Combine: samples with transparent hits
// ---------------------------------------------------------------------------
// CqImagePixel::Combine (imagepixel.cpp:144-332, no CSG) for each sample
// the sampling kernel listed: its occluding hit (already in the sample
// buffer, if any) joins its transparent hits, the list is sorted by depth as
// std::sort does, then composited back to front (a matte hit cuts out what
// lies behind it); the depth is chosen by the depth filter: min, the nearest
// hit whose opacity reaches zthreshold, else the occluding depth; midpoint,
// the mean of the two nearest such depths (the second nearest starting as
// the occluding depth); max, the farthest such depth; average, the mean
// depth of the hits with any opacity channel at zthreshold.  The front-most
// hit gives the id, and its depth goes to frontDepth.  One thread per listed
// sample, one wave per workgroup.
//
// Where the occluding hit's surface is found: its micropolygon record, of
// occ.recFloats floats with the grid index (int bits) at occ.gridField, and
// the grid's surface.  occ.mprec is null when no surface is a matte: then
// no hit's matte flag is looked at.
struct OcclRec {
    const float* mprec;
    int recFloats, gridField;
    const GridDev* grids;
};

// CqImagePixel::Combine of one listed sample (see rh_combine); d, x: room
// for its hits and its occluding hit; stk: the sort's work stack.  Inlined
// into each kernel, so that rh_combine keeps its lists in LDS with LDS
// instructions.  FULL: a depth filter other than min or some matte surface
// (without, the steps that only they take are compiled out).
template <bool FULL>
__device__ __forceinline__ void combine_sample(const FrameDev* __restrict__ fr, const uint4 e, float* d, uint32_t* x, int* stk,
                               bool longList, const float4* __restrict__ hitData,
                               const uint32_t* __restrict__ hitNext, const PrimDev* __restrict__ prims,
                               const OcclRec occ, float4* __restrict__ samples, float* __restrict__ frontDepth) {
    const size_t sidx = (size_t)e.x | ((size_t)e.y << 32);
    uint32_t head = e.z;
    int n = (int)e.w;
    const float4 oA = samples[2 * sidx], oB = samples[2 * sidx + 1];
    // the sampler's occluding depth (SqSampleData::occlZ): the occluding
    // hit's depth (or FLT_MAX without one), or with the midpoint store the
    // second nearest opaque depth, which the sampler leaves in frontDepth
    const float occlZ = FULL && fr->midpointStore ? frontDepth[sidx] : oA.w;
    // the list is linked newest first; rebuild insertion order
    for (int q = n - 1; q >= 0; --q) {
        x[q] = head;
        d[q] = hitData[head].z;
        head = hitNext[2 * (size_t)head];
    }
    const uint32_t occId = __float_as_uint(oB.w);
    bool occMatte = false;
    if (occId != 0xffffffffu) {
        x[n] = 0xffffffffu;
        d[n] = oA.w;
        ++n;
        if (FULL && occ.mprec) {
            const int g = __float_as_int(occ.mprec[(size_t)occId * occ.recFloats + occ.gridField]);
            occMatte = prims[occ.grids[g].prim].matte == RH_DEV_MATTE;
        }
    }
    if (longList) rhsort::sort_long(d, x, n, stk);
    else rhsort::sort_with(d, x, n, stk);
    float col[3] = {0.0f, 0.0f, 0.0f}, opa[3] = {0.0f, 0.0f, 0.0f};
    float opaqueDepth0 = occlZ, opaqueDepth1 = FLT_MAX, maxOpaqueDepth = FLT_MAX;
    const float zt0 = fr->zThreshold[0], zt1 = fr->zThreshold[1], zt2 = fr->zThreshold[2];
    for (int q = n - 1; q >= 0; --q) {
        float4 hA = oA, hB = oB;
        bool matte = occMatte;
        if (x[q] != 0xffffffffu) {
            // InterpolateOutputs of the hit's surface at its (u, v)
            const float4 h = hitData[x[q]];
            const PrimDev& P = prims[hitNext[2 * (size_t)x[q] + 1]];
            shade_hit(P, h.x, h.y, h.z, __float_as_uint(h.w), hA, hB);
            matte = FULL && occ.mprec && P.matte == RH_DEV_MATTE;
        }
        const float Ci[3] = {hA.x, hA.y, hA.z}, Oi[3] = {hB.x, hB.y, hB.z};
        if (FULL && matte) {
            // lerp(t, x0, 0) = (1 - t) x0 + t 0 (math.h:88-91), the colour
            // weighted by the opacity and the opacity by the colour
            for (int c = 0; c < 3; ++c) {
                col[c] = (1.0f - Oi[c]) * col[c] + Oi[c] * 0.0f;
                opa[c] = (1.0f - Ci[c]) * opa[c] + Ci[c] * 0.0f;
            }
        } else {
            for (int c = 0; c < 3; ++c) {
                const float oc = Oi[c] < 0.0f ? 0.0f : (Oi[c] > 1.0f ? 1.0f : Oi[c]);
                col[c] = col[c] * (1.0f - oc) + Ci[c];
                opa[c] = (1.0f - opa[c]) * Oi[c] + opa[c];
            }
        }
        if (Oi[0] >= zt0 && Oi[1] >= zt1 && Oi[2] >= zt2) {
            opaqueDepth1 = opaqueDepth0;
            opaqueDepth0 = hA.w;
            if (!(maxOpaqueDepth < FLT_MAX)) maxOpaqueDepth = hA.w;
        }
    }
    float depth = opaqueDepth0;
    const int df = FULL ? fr->depthFilter : 0;
    if (df == RH_DEV_DEPTH_MIDPOINT) {
        depth = n > 1 ? (opaqueDepth0 + opaqueDepth1) * 0.5f : FLT_MAX;
    } else if (df == RH_DEV_DEPTH_MAX) {
        depth = maxOpaqueDepth;
    } else if (df == RH_DEV_DEPTH_AVERAGE) {
        // front to back over the sorted list (0 / 0 without such a hit)
        float tot = 0.0f;
        int cnt = 0;
        for (int q = 0; q < n; ++q) {
            float4 hB = oB;
            float D = oA.w;
            if (x[q] != 0xffffffffu) {
                const float4 h = hitData[x[q]];
                float4 hA;
                shade_hit(prims[hitNext[2 * (size_t)x[q] + 1]], h.x, h.y, h.z, __float_as_uint(h.w), hA, hB);
                D = h.z;
            }
            if (hB.x >= zt0 || hB.y >= zt1 || hB.z >= zt2) {
                tot += D;
                ++cnt;
            }
        }
        depth = tot / (float)cnt;
    }
    const uint32_t front = x[0] == 0xffffffffu ? occId : __float_as_uint(hitData[x[0]].w);
    samples[2 * sidx] = make_float4(col[0], col[1], col[2], depth);
    frontDepth[sidx] = d[0];
    samples[2 * sidx + 1] = make_float4(opa[0], opa[1], opa[2], __uint_as_float(front));
}

#define RH_COMBINE_WG 64
template <bool FULL>
__global__ void __launch_bounds__(RH_COMBINE_WG)
rh_combine(const FrameDev* __restrict__ fr, const uint4* __restrict__ list,
                                                  uint32_t* __restrict__ hitCounter,
                                                  const float4* __restrict__ hitData,
                                                  const uint32_t* __restrict__ hitNext,
                                                  const PrimDev* __restrict__ prims, const OcclRec occ,
                                                  float4* __restrict__ samples, float* __restrict__ frontDepth) {
    __shared__ float cmbD[RH_COMBINE_WG * (RH_MAX_HITS + 1)];
    __shared__ uint32_t cmbX[RH_COMBINE_WG * (RH_MAX_HITS + 1)];
    __shared__ int cmbS[RH_COMBINE_WG * 25];
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= hitCounter[2]) return;
    const uint4 e = list[t];
    if (e.w >= RH_MAX_HITS) {
        // more hits than the LDS arrays hold: combined by rh_combine_long
        // (hitCounter[1] samples, hitCounter[3] list entries)
        atomicAdd(&hitCounter[1], 1u);
        atomicAdd(&hitCounter[3], e.w + 1u);
        return;
    }
    // the hit list and the sort's work stack in LDS (no scratch memory),
    // each lane's array an odd number of words from the next
    combine_sample<FULL>(fr, e, cmbD + threadIdx.x * (RH_MAX_HITS + 1), cmbX + threadIdx.x * (RH_MAX_HITS + 1),
                   cmbS + threadIdx.x * 25, false, hitData, hitNext, prims, occ, samples, frontDepth);
}

// The samples rh_combine left (more than RH_MAX_HITS - 1 transparent hits,
// as many as the reference's unbounded lists hold): the same steps with the
// list and the sort's stack in global work arrays (the sort restated for
// any length), one thread per listed sample.  workD and workX hold
// hitCounter[3] entries, workStk 192 ints per deferred sample
// (hitCounter[1] of them); hitCounter[4] and [5] hand them out.
template <bool FULL>
__global__ void __launch_bounds__(64)
rh_combine_long(const FrameDev* __restrict__ fr, const uint4* __restrict__ list, uint32_t* __restrict__ hitCounter,
                const float4* __restrict__ hitData, const uint32_t* __restrict__ hitNext,
                const PrimDev* __restrict__ prims, const OcclRec occ, float4* __restrict__ samples,
                float* __restrict__ frontDepth, float* __restrict__ workD, uint32_t* __restrict__ workX,
                int* __restrict__ workStk) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= hitCounter[2]) return;
    const uint4 e = list[t];
    if (e.w < RH_MAX_HITS) return;
    const uint32_t off = atomicAdd(&hitCounter[4], e.w + 1u);
    const uint32_t k = atomicAdd(&hitCounter[5], 1u);
    combine_sample<FULL>(fr, e, workD + off, workX + off, workStk + (size_t)k * 192, true, hitData, hitNext, prims, occ,
                   samples, frontDepth);
}


// ---------------------------------------------------------------------------
// Filter + expose + quantise
// ---------------------------------------------------------------------------
// Normalise one pixel's weighted sums and write its channels and 8-bit rgba:
// FilterBucket's tail (bucketprocessor.cpp:640-707), ExposeBucket
// (bucketprocessor.cpp:766-806) and FormatBucketForDisplay
// (ddmanager.cpp:1045-1118).  acc = weighted Ci.rgb, Oi.rgb, z.
__device__ __forceinline__ void finish_pixel(const FrameDev* __restrict__ fr, const float* acc, float gTot, int count,
                                             int ns, size_t t, const float* __restrict__ dither,
                                             float* __restrict__ channels, uchar4* __restrict__ rgba8) {
    float ch[9];
    float coverage;
    if (count == 0) {
        for (int k = 0; k < 9; ++k) ch[k] = 0.0f;
        ch[7] = FLT_MAX;
        coverage = 0.0f;
    } else {
        float inv = (float)(1.0 / (double)gTot);
        for (int k = 0; k < 6; ++k) ch[k] = acc[k] * inv;
        ch[7] = acc[6] * inv;
        coverage = (count >= ns) ? 1.0f : (float)count / (float)ns;
    }
    float a = (ch[3] + ch[4] + ch[5]) / 3.0f;
    ch[6] = a * coverage;
    ch[8] = coverage;
    const float gain = fr->exposureGain, gamma = fr->exposureGamma;
    if (count > 0 && !(gain == 1.0f && gamma == 1.0f)) {
        float oog = 1.0f / gamma;
        for (int k = 0; k < 3; ++k) {
            if (gain != 1.0f) ch[k] *= gain;
            if (gamma != 1.0f) ch[k] = (float)pow((double)ch[k], (double)oog);
        }
    }
    float* out = channels + t * 9;
    for (int k = 0; k < 9; ++k) out[k] = ch[k];
    const double s = (double)dither[t];
    const float vals[4] = {ch[0], ch[1], ch[2], ch[6]};
    unsigned char q[4];
    for (int k = 0; k < 4; ++k) {
        double value = vals[k];
        if (fr->qOne != 0.0f) {
            double xq = 0.0f + value * (double)(fr->qOne - 0.0f) + ((double)fr->qDither * s);
            value = (double)(lfloor_d(xq - 0.5) + 1);
            value = value < (double)fr->qMin ? (double)fr->qMin : (value > (double)fr->qMax ? (double)fr->qMax : value);
        }
        q[k] = (unsigned char)value;
    }
    rgba8[t] = make_uchar4(q[0], q[1], q[2], q[3]);
}

// General filter: one thread per display pixel, samples read straight from
// HBM.  Used when a tile's halo does not fit the tiled kernel's LDS.
// FilterBucket, non-separable branch (bucketprocessor.cpp:561-663).
__global__ void __launch_bounds__(256) rh_filter(const FrameDev* __restrict__ fr, const float4* __restrict__ samples,
                                                 const uchar4* __restrict__ pixPattern, const float* __restrict__ pos2d,
                                                 const float* __restrict__ filterTab, const float* __restrict__ dither,
                                                 float* __restrict__ channels, uchar4* __restrict__ rgba8) {
    const int W = fr->cropX1 - fr->cropX0, H = fr->cropY1 - fr->cropY0;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= W * H) return;
    const int x = fr->cropX0 + t % W, y = fr->cropY0 + t / W;
    const int ns = fr->ns, xm = fr->shiftX, ym = fr->shiftY;
    const float xf = fr->xfwo2, yf = fr->yfwo2;
    const float xcent = x + 0.5f, ycent = y + 0.5f;
    float gTot = 0.0f;
    int count = 0;
    float acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int fy = -ym; fy <= ym; ++fy)
        for (int fx = -xm; fx <= xm; ++fx) {
            const int qx = x + fx, qy = y + fy;
            const size_t q = (size_t)(qy - fr->extY0) * fr->extW + (qx - fr->extX0);
            const int pat = pixPattern[q].x;
            const float* wt = filterTab + ((fy + ym) * (2 * xm + 1) + fx + xm) * ns;
            for (int si = 0; si < ns; ++si) {
                float vx = ((float)qx + pos2d[2 * (pat * ns + si)]) - xcent;
                float vy = ((float)qy + pos2d[2 * (pat * ns + si) + 1]) - ycent;
                if (vx >= -xf && vy >= -yf && vx <= xf && vy <= yf) {
                    float g = wt[si];
                    gTot += g;
                    const float4 a = samples[2 * (q * ns + si)], b = samples[2 * (q * ns + si) + 1];
                    if (__float_as_uint(b.w) != 0xffffffffu) {
                        acc[0] += a.x * g; acc[1] += a.y * g; acc[2] += a.z * g;
                        acc[3] += b.x * g; acc[4] += b.y * g; acc[5] += b.z * g;
                        acc[6] += a.w * g;
                        ++count;
                    }
                }
            }
        }
    finish_pixel(fr, acc, gTot, count, ns, (size_t)t, dither, channels, rgba8);
}

// Box-test masks of every sample of the extended region, once per frame
// shape: bit (fx + xm) says that the sample of pixel qx lies inside the
// filter box of pixel qx - fx (|position - pixel centre| <= ceil(width)/2,
// bucketprocessor.cpp:600-612, evaluated exactly as the filter loop does),
// bit (8 + fy + ym) the same for rows.
__global__ void __launch_bounds__(256) rh_box_masks(const FrameDev* __restrict__ fr,
                                                    const uchar4* __restrict__ pixPattern,
                                                    const float* __restrict__ pos2d, uint16_t* __restrict__ masks) {
    const int ns = fr->ns, xm = fr->shiftX, ym = fr->shiftY, extW = fr->extW;
    const size_t n = (size_t)extW * fr->extH * ns;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const size_t q = t / ns;
    const int si = (int)(t - q * ns);
    const int qx = fr->extX0 + (int)(q % extW), qy = fr->extY0 + (int)(q / extW);
    const int pat = pixPattern[q].x;
    const float sx = (float)qx + pos2d[2 * (pat * ns + si)];
    const float sy = (float)qy + pos2d[2 * (pat * ns + si) + 1];
    const float xf = fr->xfwo2, yf = fr->yfwo2;
    uint32_t mask = 0;
    for (int fx = -xm; fx <= xm; ++fx) {
        const float xcent = (qx - fx) + 0.5f;
        const float vx = sx - xcent;
        if (vx >= -xf && vx <= xf) mask |= 1u << (fx + xm);
    }
    for (int fy = -ym; fy <= ym; ++fy) {
        const float ycent = (qy - fy) + 0.5f;
        const float vy = sy - ycent;
        if (vy >= -yf && vy <= yf) mask |= 1u << (8 + fy + ym);
    }
    masks[t] = (uint16_t)mask;
}

// Tiled filter.  One workgroup per RH_FT x RH_FT tile of display pixels,
// four waves over the channels: wave 0 sums Ci.r and Ci.g (with the weight
// total and the count), wave 1 Ci.b and z, wave 2 Oi.r and Oi.g, wave 3
// Oi.b; each lane one pixel, each sum in the reference's order (filter rows,
// filter columns, samples), so the split changes no rounding.  The tile's halo of samples is
// staged in LDS once, with RH_FU coalesced loads in flight per lane, as 8
// dwords per sample: Ci.rgb, z, Oi.rgb and a word holding the sample's
// box-test mask (bits 0..15) and whether it holds a surface (bit 31).  Halo
// pixel records are 8*ns+RH_FPAD dwords apart so that the lanes of a wave read
// different LDS banks.
#define RH_FT 8
#define RH_FU 16
#define RH_FTHREADS 256
#ifndef RH_FPAD
#define RH_FPAD 1
#endif

// The weighted sums of one pixel over its filter footprint, in the
// reference's order (filter rows, filter columns, samples): NC channels
// starting at `rec0` of each staged sample record (8 words apart), the
// sample's mask word at offset `mw`.  A sample outside the pixel's box adds
// nothing, one inside adds its weight to gTot, and one holding a surface also
// adds weight x value to each channel sum (gTot and count only with TOT).
// The adds are unconditional with the weight zeroed where the reference
// skips: x + 0 * v == x exactly, since the sums start at +0 and never become
// -0.  Weights are read with uniform (scalar) loads.
template <int NC, bool TOT>
__device__ __forceinline__ void filter_taps(const float* __restrict__ rec0, const float* __restrict__ filterTab,
                                            int ns, int xm, int ym, int HX, int S, int lx, int ly, int mw,
                                            float* acc, float& gTot, int& count) {
    for (int fy = -ym; fy <= ym; ++fy)
        for (int fx = -xm; fx <= xm; ++fx) {
            const float* rec = rec0 + ((ly + fy + ym) * HX + (lx + fx + xm)) * S;
            const float* wt = filterTab + ((fy + ym) * (2 * xm + 1) + fx + xm) * ns;
            const uint32_t need = (1u << (fx + xm)) | (1u << (8 + fy + ym));
            // four samples at a time, every LDS read issued first
            int si = 0;
            for (; si + 4 <= ns; si += 4) {
                uint32_t m[4];
                float d[4][NC];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float* r = rec + (si + q) * 8;
                    m[q] = __float_as_uint(r[mw]);
#pragma unroll
                    for (int c = 0; c < NC; ++c) d[q][c] = r[c];
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const bool in = (m[q] & need) == need;
                    const bool hit = in && (m[q] >> 31);
                    const float g = wt[si + q];
                    const float ghit = hit ? g : 0.0f;
                    if (TOT) {
                        gTot = gTot + (in ? g : 0.0f);
                        count += hit ? 1 : 0;
                    }
#pragma unroll
                    for (int c = 0; c < NC; ++c) acc[c] = acc[c] + d[q][c] * ghit;
                }
            }
            for (; si < ns; ++si) {
                const float* r = rec + si * 8;
                const uint32_t mm = __float_as_uint(r[mw]);
                const bool in = (mm & need) == need;
                const bool hit = in && (mm >> 31);
                const float g = wt[si];
                const float ghit = hit ? g : 0.0f;
                if (TOT) {
                    gTot = gTot + (in ? g : 0.0f);
                    count += hit ? 1 : 0;
                }
#pragma unroll
                for (int c = 0; c < NC; ++c) acc[c] = acc[c] + r[c] * ghit;
            }
        }
}
__global__ void __launch_bounds__(RH_FTHREADS) rh_filter_tiled(const FrameDev* __restrict__ fr,
                                                       const float4* __restrict__ samples,
                                                       const uint16_t* __restrict__ boxMask,
                                                       const float* __restrict__ filterTab,
                                                       const float* __restrict__ dither, float* __restrict__ channels,
                                                       uchar4* __restrict__ rgba8, int tilesX, int ntiles) {
    // LDS: halo records (RH_FT+2*xm)*(RH_FT+2*ym)*(8*ns+RH_FPAD) floats; the
    // tail exchange reuses the first 320 floats.
    extern __shared__ float lds[];
    // Workgroups are dealt round-robin to the 8 XCDs; give each XCD a
    // contiguous run of tiles so neighbouring tiles share halos in its L2.
    // The grid is a multiple of 8 workgroups.
    const int per = gridDim.x >> 3;
    const int tile = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (tile >= ntiles) return;
    const int ns = fr->ns, xm = fr->shiftX, ym = fr->shiftY;
    const int extX0 = fr->extX0, extY0 = fr->extY0, extW = fr->extW, extH = fr->extH;
    const int tx0 = fr->cropX0 + (tile % tilesX) * RH_FT, ty0 = fr->cropY0 + (tile / tilesX) * RH_FT;
    const int HX = RH_FT + 2 * xm, HY = RH_FT + 2 * ym, NH = HX * HY;
    const int S = 8 * ns + RH_FPAD;
    const int tid = threadIdx.x;

    // 1. Stage the halo's sample records (2 float4 per sample) and masks.
    {
        const int per_px = 2 * ns, total = NH * per_px;
        const int dhp = RH_FTHREADS / per_px, dpart = RH_FTHREADS % per_px;
        int hp = tid / per_px, part = tid % per_px;
        int hy = hp / HX, hx = hp % HX;
        for (int base = 0; base < total; base += RH_FTHREADS * RH_FU) {
            float4 v[RH_FU];
            uint32_t mk[RH_FU];
            int dst[RH_FU];
#pragma unroll
            for (int u = 0; u < RH_FU; ++u) {
                const int i = base + u * RH_FTHREADS + tid;
                v[u] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0xffffffffu));
                mk[u] = 0;
                dst[u] = -1;
                if (i < total) {
                    const int qx = tx0 - xm + hx, qy = ty0 - ym + hy;
                    dst[u] = (hp * S + part * 4) * 2 + (part & 1);
                    if (qx >= extX0 && qx < extX0 + extW && qy >= extY0 && qy < extY0 + extH) {
                        const size_t q = (size_t)(qy - extY0) * extW + (qx - extX0);
                        v[u] = samples[q * per_px + part];
                        if (part & 1) mk[u] = boxMask[q * ns + (part >> 1)];
                    }
                }
                part += dpart;
                int adv = dhp;
                if (part >= per_px) { part -= per_px; ++adv; }
                hp += adv;
                hx += adv;
                while (hx >= HX) { hx -= HX; ++hy; }
            }
#pragma unroll
            for (int u = 0; u < RH_FU; ++u) {
                if (dst[u] < 0) continue;
                float* d = lds + (dst[u] >> 1);
                d[0] = v[u].x;
                d[1] = v[u].y;
                d[2] = v[u].z;
                d[3] = (dst[u] & 1) ? __uint_as_float(mk[u] | ((__float_as_uint(v[u].w) != 0xffffffffu) ? 0x80000000u : 0u))
                                    : v[u].w;
            }
        }
    }
    __syncthreads();
    // 2. Weighted sums, one lane per pixel per wave's channels.
    const int lane = tid & 63, part = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lx = lane & (RH_FT - 1), ly = lane / RH_FT;
    const int x = tx0 + lx, y = ty0 + ly;
    const bool valid = x < fr->cropX1 && y < fr->cropY1;
    float gTot = 0.0f;
    int count = 0;
    float acc[2] = {0.0f, 0.0f};
    if (valid) {
        // record words: Ci.rgb, z, Oi.rgb, mask
        if (part == 0) filter_taps<2, true>(lds, filterTab, ns, xm, ym, HX, S, lx, ly, 7, acc, gTot, count);
        else if (part == 1) filter_taps<2, false>(lds + 2, filterTab, ns, xm, ym, HX, S, lx, ly, 5, acc, gTot, count);
        else if (part == 2) filter_taps<2, false>(lds + 4, filterTab, ns, xm, ym, HX, S, lx, ly, 3, acc, gTot, count);
        else filter_taps<1, false>(lds + 6, filterTab, ns, xm, ym, HX, S, lx, ly, 1, acc, gTot, count);
    }
    __syncthreads();                  // the halo records are no longer read
    // waves 1..3 hand their sums to wave 0: Ci.b, z, Oi.r, Oi.g, Oi.b
    if (part == 1) {
        lds[lane * 5 + 0] = acc[0];
        lds[lane * 5 + 1] = acc[1];
    } else if (part == 2) {
        lds[lane * 5 + 2] = acc[0];
        lds[lane * 5 + 3] = acc[1];
    } else if (part == 3) {
        lds[lane * 5 + 4] = acc[0];
    }
    __syncthreads();
    if (part == 0 && valid) {
        const float full[7] = {acc[0], acc[1], lds[lane * 5 + 0], lds[lane * 5 + 2], lds[lane * 5 + 3],
                               lds[lane * 5 + 4], lds[lane * 5 + 1]};
        const size_t t = (size_t)(y - fr->cropY0) * (fr->cropX1 - fr->cropX0) + (x - fr->cropX0);
        finish_pixel(fr, full, gTot, count, ns, t, dither, channels, rgba8);
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Occlusion culling: each bucket's tree of sub-sample depths
// (CqOcclusionTree, libs/core/occlusion.cpp) and the cull tests of every
// (surface, bucket) pair of a scene, on the depths of a frame.
// ---------------------------------------------------------------------------
// The opaque depth (occlZ) of sub-sample cell (cx, cy) of a bucket's sample
// region, from a frame's sample records (the samplers store it in the first
// record's w before any combine).
__device__ __forceinline__ float cell_depth(const FrameDev* __restrict__ fr, const CullTreeDev& t,
                                            const float4* __restrict__ samples, int cx, int cy) {
    const int nx = fr->nx, ny = fr->ny;
    const int px = t.sx0 + cx / nx, py = t.sy0 + cy / ny;
    const int si = (cy - (cy / ny) * ny) * nx + (cx - (cx / nx) * nx);
    const size_t sidx = ((size_t)(py - fr->extY0) * fr->extW + (px - fr->extX0)) * fr->ns + si;
    return samples[2 * sidx].w;
}

// One workgroup per bucket: its depth pyramid, level 0 the cells, each
// further level the maximum of 2x2 cells of the one below.
__global__ void __launch_bounds__(256) rh_cull_pyramid(const FrameDev* __restrict__ fr,
                                                       const CullTreeDev* __restrict__ trees,
                                                       const float4* __restrict__ samples, float* __restrict__ pyr) {
    const CullTreeDev t = trees[blockIdx.x];
    if (t.levels == 0) return;
    float* L = pyr + t.pyr;
    const int n0 = t.cw * t.ch;
    for (int i = threadIdx.x; i < n0; i += blockDim.x) L[i] = cell_depth(fr, t, samples, i % t.cw, i / t.cw);
    int w = t.cw, h = t.ch;
    for (int l = 1; l < t.levels; ++l) {
        __syncthreads();
        float* U = L + (size_t)w * h;
        const int uw = cull_level_dim(t.cw, l), uh = cull_level_dim(t.ch, l);
        for (int i = threadIdx.x; i < uw * uh; i += blockDim.x) {
            const int x = i % uw, y = i / uw;
            const int x0 = 2 * x, y0 = 2 * y;
            float m = L[(size_t)y0 * w + x0];
            if (x0 + 1 < w) m = fmaxf(m, L[(size_t)y0 * w + x0 + 1]);
            if (y0 + 1 < h) {
                m = fmaxf(m, L[(size_t)(y0 + 1) * w + x0]);
                if (x0 + 1 < w) m = fmaxf(m, L[(size_t)(y0 + 1) * w + x0 + 1]);
            }
            U[i] = m;
        }
        L = U;
        w = uw;
        h = uh;
    }
}

// One thread per (surface, bucket) pair: the largest depth among the tree
// leaves the surface's raster bound meets (padding leaves count as 0), or
// -1 when it meets none; the surface is culled there when this lies below
// its minimum depth (canCull, occlusion.cpp:143-196).
__global__ void __launch_bounds__(256) rh_cull_pairs(const CullTreeDev* __restrict__ trees, int bx0, int by0, int nbr,
                                                     const float4* __restrict__ bounds,
                                                     const int4* __restrict__ ranges,
                                                     const unsigned long long* __restrict__ start, int nnodes,
                                                     unsigned long long npairs, const float* __restrict__ pyr,
                                                     float* __restrict__ pairMax) {
    const unsigned long long pi = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (pi >= npairs) return;
    // the surface: the last one whose first pair is at or before pi
    int lo = 0, hi = nnodes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= pi) lo = mid;
        else hi = mid - 1;
    }
    const int4 R = ranges[lo];
    const float4 bd = bounds[lo];
    const unsigned long long off = pi - start[lo];
    const int nbx = R.z - R.x + 1;
    const int col = R.x + (int)(off % (unsigned long long)nbx), row = R.y + (int)(off / (unsigned long long)nbx);
    const CullTreeDev t = trees[(row - by0) * nbr + (col - bx0)];
    int cx0 = 0, cx1 = -1, cy0 = 0, cy1 = -1;
    const int kind = t.levels ? cull_cells(t, bd.x, bd.y, bd.z, bd.w, cx0, cx1, cy0, cy1) : 0;
    float m = -1.0f;
    if (kind == 2) m = 0.0f;
    if (kind != 0 && cx0 <= cx1 && cy0 <= cy1) m = fmaxf(m, pyramid_max(pyr + t.pyr, t.cw, t.ch, cx0, cx1, cy0, cy1));
    pairMax[pi] = m;
}

// One workgroup per probe: the largest depth over a probe's cells, from a
// frame's sample records rendered up to the probe's moment.
// Several probes at once: probe i reads the leaves of bucket pb[i].x, kind
// pb[i].y, cells pc[i] (cx0, cx1, cy0, cy1), into out[pb[i].z].
__global__ void __launch_bounds__(256) rh_cull_probes_max(const FrameDev* __restrict__ fr,
                                                          const CullTreeDev* __restrict__ trees,
                                                          const int4* __restrict__ pb, const int4* __restrict__ pc,
                                                          const float4* __restrict__ samples,
                                                          float* __restrict__ out) {
    __shared__ float part[4];
    const int4 b = pb[blockIdx.x], c = pc[blockIdx.x];
    const int kind = b.y;
    const CullTreeDev t = trees[b.x];
    float m = -FLT_MAX;
    const int w = c.y - c.x + 1, h = c.w - c.z + 1;
    if (kind != 0 && w > 0 && h > 0)
        for (int i = threadIdx.x; i < w * h; i += blockDim.x)
            m = fmaxf(m, cell_depth(fr, t, samples, c.x + i % w, c.z + i / w));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = fmaxf(fmaxf(part[0], part[1]), fmaxf(part[2], part[3]));
        if (kind == 2) r = fmaxf(r, 0.0f);
        if (kind == 0 || r == -FLT_MAX) r = kind == 2 ? 0.0f : -1.0f;
        out[b.z] = r;
    }
}

extern "C" hipError_t rh_launch_cull_probes_max(hipStream_t st, const FrameDev* fr, const CullTreeDev* trees,
                                                const int4* pb, const int4* pc, int n, const float4* samples,
                                                float* out) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(rh_cull_probes_max, dim3(n), dim3(256), 0, st, fr, trees, pb, pc, samples, out);
    return hipGetLastError();
}

extern "C" hipError_t rh_launch_cull_pyramid(hipStream_t st, const FrameDev* fr, const CullTreeDev* trees, int nbuckets,
                                             const float4* samples, float* pyr) {
    if (nbuckets <= 0) return hipSuccess;
    hipLaunchKernelGGL(rh_cull_pyramid, dim3(nbuckets), dim3(256), 0, st, fr, trees, samples, pyr);
    return hipGetLastError();
}

extern "C" hipError_t rh_launch_cull_pairs(hipStream_t st, const CullTreeDev* trees, int bx0, int by0, int nbr,
                                           const float4* bounds, const int4* ranges, const unsigned long long* start,
                                           int nnodes, unsigned long long npairs, const float* pyr, float* pairMax) {
    if (npairs == 0) return hipSuccess;
    const unsigned long long blocks = (npairs + 255) / 256;
    if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(rh_cull_pairs, dim3((unsigned)blocks), dim3(256), 0, st, trees, bx0, by0, nbr, bounds, ranges,
                       start, nnodes, npairs, pyr, pairMax);
    return hipGetLastError();
}

extern "C" hipError_t rh_launch_dice(hipStream_t st, const FrameDev* fr, const GridDev* grids, int ngrids,
                                     const PrimDev* prims, const float* params, float* mprec,
                                     float4* mpbound, float4* mprecMB) {
    if (ngrids <= 0) return hipSuccess;
    hipLaunchKernelGGL(rh_dice_bust, dim3(ngrids), dim3(256), 0, st, fr, grids, prims, params, mprec,
                       mpbound, mprecMB);
    return hipGetLastError();
}

template <int MT, bool POW2, bool MB>
static void launch_bin_t(hipStream_t st, size_t wgs, const FrameDev* fr, const BucketDev* buckets, int nbuckets,
                         const TaskDev* tasks, const GridDev* grids, const uint32_t* candStart,
                         const uint32_t* candGrid, const uint32_t* segBucket, size_t nsegments,
                         const unsigned long long* capStart, const float4* mpbound, const float4* mprecMB,
                         uint32_t* segCounts, uint32_t* lists, uint32_t* listCount, uint32_t* bucketRefs,
                         unsigned long long* counters) {
    if (nsegments > 0)
        hipLaunchKernelGGL((rh_bin_count<MT, POW2, MB>), dim3((unsigned)wgs), dim3(256), 0, st, fr, buckets, tasks,
                           grids, candGrid, segBucket, nsegments, mpbound, mprecMB, segCounts);
    hipLaunchKernelGGL(rh_bin_scan, dim3(nbuckets), dim3(256), 0, st, buckets, candStart, capStart, segCounts,
                       listCount, bucketRefs, counters);
    if (nsegments > 0)
        hipLaunchKernelGGL((rh_bin_write<MT, POW2, MB>), dim3((unsigned)wgs), dim3(256), 0, st, fr, buckets, tasks,
                           grids, candGrid, segBucket, nsegments, mpbound, mprecMB, segCounts, capStart, lists);
}

template <bool MB>
static void launch_bin_mb(hipStream_t st, size_t wgs, int taskSide, int pow2, const FrameDev* fr,
                          const BucketDev* buckets, int nbuckets, const TaskDev* tasks, const GridDev* grids,
                          const uint32_t* candStart, const uint32_t* candGrid, const uint32_t* segBucket,
                          size_t nsegments, const unsigned long long* capStart, const float4* mpbound,
                          const float4* mprecMB, uint32_t* segCounts, uint32_t* lists, uint32_t* listCount,
                          uint32_t* bucketRefs, unsigned long long* counters) {
    auto* f = taskSide <= 2 ? (pow2 ? &launch_bin_t<2, true, MB> : &launch_bin_t<2, false, MB>)
                            : (pow2 ? &launch_bin_t<4, true, MB> : &launch_bin_t<4, false, MB>);
    f(st, wgs, fr, buckets, nbuckets, tasks, grids, candStart, candGrid, segBucket, nsegments, capStart, mpbound,
      mprecMB, segCounts, lists, listCount, bucketRefs, counters);
}

// taskSide: the largest task grid side of any bucket (at most 4); pow2: both
// bucket sizes are powers of two; mprecMB: the motion records of a frame
// with moving micropolygons (their pixel windows), else null.
extern "C" hipError_t rh_launch_bin(hipStream_t st, const FrameDev* fr, const BucketDev* buckets, int nbuckets,
                                    const TaskDev* tasks, const GridDev* grids, const uint32_t* candStart,
                                    const uint32_t* candGrid, const uint32_t* segBucket, size_t nsegments,
                                    const unsigned long long* capStart, const float4* mpbound, uint32_t* segCounts,
                                    uint32_t* lists, uint32_t* listCount, uint32_t* bucketRefs,
                                    unsigned long long* counters, int taskSide, int pow2, const float4* mprecMB) {
    if (nbuckets <= 0) return hipSuccess;
    if (taskSide < 1 || taskSide > 4) return hipErrorInvalidValue;
    // enough waves to fill the device; each walks segments grid-stride
    const size_t wgs = std::min<size_t>((nsegments + 3) / 4, 8192);
    if (mprecMB)
        launch_bin_mb<true>(st, wgs, taskSide, pow2, fr, buckets, nbuckets, tasks, grids, candStart, candGrid,
                            segBucket, nsegments, capStart, mpbound, mprecMB, segCounts, lists, listCount, bucketRefs,
                            counters);
    else
        launch_bin_mb<false>(st, wgs, taskSide, pow2, fr, buckets, nbuckets, tasks, grids, candStart, candGrid,
                             segBucket, nsegments, capStart, mpbound, nullptr, segCounts, lists, listCount,
                             bucketRefs, counters);
    return hipGetLastError();
}

template <int CH, int SLOTS, int STORE = RH_STORE_OCCLUDE, int WPE = 4>
static void launch_sample_t(hipStream_t st, int grid, const SampleLaunch& a) {
    hipLaunchKernelGGL((rh_sample<CH, SLOTS, STORE, WPE>), dim3(grid), dim3(256), 0, st, a.fr, a.buckets, a.tasks, a.items,
                       a.itemBase, a.nStatic, a.nDevice, a.persistent, a.queueHead, a.capStart, a.listCount, a.lists,
                       reinterpret_cast<const float4*>(a.mprec), a.grids, a.prims, a.pos2d, a.pos1d, a.pixPattern,
                       a.samples, a.hitData, a.hitNext, a.hitCounter, a.combineList, a.frontDepth, a.rowCounts,
                       a.rowBase, a.idCut, a.itemCut, a.depthOnly);
}

// One sampling launch of `grid` workgroups with kernel variant `variant`
// (RH_VARIANT_CH / RH_VARIANT_SLOTS).  The midpoint and list-all stores
// take one form each (96-record chunks, 2 blocks per wave) whatever the
// variant: a wave loops over an item's blocks, so any form covers any item.
extern "C" hipError_t rh_launch_sample(hipStream_t st, int variant, int grid, const SampleLaunch* a) {
    if (grid <= 0) return hipSuccess;
    if (variant < 0 || variant >= RH_SAMPLE_VARIANTS) return hipErrorInvalidValue;
    if (a->store == RH_STORE_MIDPOINT) {
        launch_sample_t<96, 2, RH_STORE_MIDPOINT>(st, grid, *a);
        return hipGetLastError();
    }
    if (a->store == RH_STORE_LIST_ALL) {
        launch_sample_t<96, 2, RH_STORE_LIST_ALL>(st, grid, *a);
        return hipGetLastError();
    }
    switch (variant) {
    case 0: launch_sample_t<216, 4>(st, grid, *a); break;
    case 1: launch_sample_t<216, 2>(st, grid, *a); break;
    case 2: launch_sample_t<96, 4>(st, grid, *a); break;
    case 3: launch_sample_t<96, 2, RH_STORE_OCCLUDE, RH_WPE_96_2>(st, grid, *a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// The motion form of the sampler: one workgroup of RH_MB_WAVES waves per
// task, tasks [a->itemBase, a->itemBase + ntasks).
#ifndef RH_MB_WAVES
#define RH_MB_WAVES 16
#endif
extern "C" hipError_t rh_launch_sample_mb(hipStream_t st, const SampleLaunch* a, int ntasks,
                                          const float4* mprecMB) {
    if (ntasks <= 0) return hipSuccess;
#define RH_LAUNCH_MB(STORE)                                                                                          \
    hipLaunchKernelGGL((rh_sample_mb<RH_MB_WAVES, STORE>), dim3(ntasks), dim3(RH_MB_WAVES * 64), 0, st, a->fr,        \
                       a->buckets, a->tasks, a->itemBase + ntasks, a->capStart, a->listCount, a->lists, mprecMB,       \
                       a->grids, a->prims, a->pos2d, a->pos1d, a->pixPattern, a->samples, a->hitData, a->hitNext,     \
                       a->hitCounter, a->combineList, a->frontDepth, a->rowCounts, a->rowBase, a->itemBase, a->idCut, \
                       a->itemCut, a->taskList, a->depthOnly)
    switch (a->store) {
    case RH_STORE_OCCLUDE: RH_LAUNCH_MB(RH_STORE_OCCLUDE); break;
    case RH_STORE_MIDPOINT: RH_LAUNCH_MB(RH_STORE_MIDPOINT); break;
    case RH_STORE_LIST_ALL: RH_LAUNCH_MB(RH_STORE_LIST_ALL); break;
    default: return hipErrorInvalidValue;
    }
#undef RH_LAUNCH_MB
    return hipGetLastError();
}

// The depth-of-field form of the sampler: one workgroup per task, tasks
// [a->itemBase, a->itemBase + ntasks).
extern "C" hipError_t rh_launch_sample_dof(hipStream_t st, const SampleLaunch* a, int ntasks, const DofTables* t) {
    if (ntasks <= 0) return hipSuccess;
#define RH_LAUNCH_DOF(STORE)                                                                                       \
    hipLaunchKernelGGL((rh_sample_dof<STORE>), dim3(ntasks), dim3(RH_DOF_WAVES * 64), 0, st, a->fr, a->buckets,     \
                       a->tasks, a->itemBase, a->capStart, a->listCount, a->lists,                                   \
                       reinterpret_cast<const float4*>(a->mprec), a->grids, a->prims, a->pos2d, a->pos1d,           \
                       a->pixPattern, t->shuffled, t->lens, t->pixDof, t->cells, a->samples, a->hitData, a->hitNext, \
                       a->hitCounter, a->combineList, a->frontDepth, a->rowCounts, a->rowBase, a->idCut, a->itemCut, \
                       a->taskList, a->depthOnly)
    switch (a->store) {
    case RH_STORE_OCCLUDE: RH_LAUNCH_DOF(RH_STORE_OCCLUDE); break;
    case RH_STORE_MIDPOINT: RH_LAUNCH_DOF(RH_STORE_MIDPOINT); break;
    case RH_STORE_LIST_ALL: RH_LAUNCH_DOF(RH_STORE_LIST_ALL); break;
    default: return hipErrorInvalidValue;
    }
#undef RH_LAUNCH_DOF
    return hipGetLastError();
}

// Workgroups of one sampling variant that fit the device at once.
extern "C" hipError_t rh_sample_occupancy(int variant, int* blocksPerCU) {
    switch (variant) {
    case 0: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, rh_sample<216, 4>, 256, 0);
    case 1: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, rh_sample<216, 2>, 256, 0);
    case 2: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, rh_sample<96, 4>, 256, 0);
    case 3:
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, rh_sample<96, 2, RH_STORE_OCCLUDE, RH_WPE_96_2>,
                                                            256, 0);
    default: return hipErrorInvalidValue;
    }
}

extern "C" hipError_t rh_launch_plan_items(hipStream_t st, const FrameDev* fr, const TaskDev* tasks, int ntasks,
                                           const uint32_t* listCount, uint32_t splitCost, int itemBlocks,
                                           uint2* items, uint32_t* itemCount, uint32_t* queueHead) {
    hipLaunchKernelGGL(rh_plan_items, dim3(1), dim3(1024), 0, st, fr, tasks, ntasks, listCount, splitCost,
                       itemBlocks, items, itemCount, queueHead);
    return hipGetLastError();
}

// maxEntries: the combine list's capacity (the extended region's samples);
// the kernel reads the actual count from hitCounter[2].
extern "C" hipError_t rh_launch_combine(hipStream_t st, const FrameDev* fr, size_t maxEntries, const uint4* list,
                                        uint32_t* hitCounter, const float4* hitData, const uint32_t* hitNext,
                                        const PrimDev* prims, const float* mprec, int recFloats, int gridField,
                                        const GridDev* grids, int full, float4* samples, float* frontDepth) {
    if (maxEntries == 0) return hipSuccess;
    const OcclRec occ = {mprec, recFloats, gridField, grids};
    const dim3 grid((unsigned)((maxEntries + RH_COMBINE_WG - 1) / RH_COMBINE_WG));
    if (full)
        hipLaunchKernelGGL(rh_combine<true>, grid, dim3(RH_COMBINE_WG), 0, st, fr, list, hitCounter, hitData, hitNext,
                           prims, occ, samples, frontDepth);
    else
        hipLaunchKernelGGL(rh_combine<false>, grid, dim3(RH_COMBINE_WG), 0, st, fr, list, hitCounter, hitData, hitNext,
                           prims, occ, samples, frontDepth);
    return hipGetLastError();
}

// The long-list pass over the combine list (up to maxEntries entries; the
// kernel reads the count from hitCounter[2]).
extern "C" hipError_t rh_launch_combine_long(hipStream_t st, const FrameDev* fr, size_t maxEntries, const uint4* list,
                                             uint32_t* hitCounter, const float4* hitData, const uint32_t* hitNext,
                                             const PrimDev* prims, const float* mprec, int recFloats, int gridField,
                                             const GridDev* grids, int full, float4* samples, float* frontDepth,
                                             float* workD, uint32_t* workX, int* workStk) {
    if (maxEntries == 0) return hipSuccess;
    const OcclRec occ = {mprec, recFloats, gridField, grids};
    const dim3 grid((unsigned)((maxEntries + 63) / 64));
    if (full)
        hipLaunchKernelGGL(rh_combine_long<true>, grid, dim3(64), 0, st, fr, list, hitCounter, hitData, hitNext, prims,
                           occ, samples, frontDepth, workD, workX, workStk);
    else
        hipLaunchKernelGGL(rh_combine_long<false>, grid, dim3(64), 0, st, fr, list, hitCounter, hitData, hitNext, prims,
                           occ, samples, frontDepth, workD, workX, workStk);
    return hipGetLastError();
}

extern "C" hipError_t rh_launch_frame_begin(hipStream_t st, unsigned long long* counters, uint32_t* hitCtl) {
    hipLaunchKernelGGL(rh_frame_begin, dim3(1), dim3(64), 0, st, counters, hitCtl);
    return hipGetLastError();
}

extern "C" hipError_t rh_launch_reduce_counters(hipStream_t st, const uint32_t* taskCounts, int ntasks,
                                                const uint32_t* bucketRefs, int nbuckets,
                                                unsigned long long* counters) {
    hipLaunchKernelGGL(rh_reduce_counters, dim3(1), dim3(1024), 0, st, taskCounts, ntasks, bucketRefs, nbuckets,
                       counters);
    return hipGetLastError();
}
// LDS bytes of one tiled-filter workgroup, or 0 if the tiled kernel cannot
// take this frame shape (mask bits or LDS size exceeded).
extern "C" size_t rh_filter_tile_lds(int ns, int xm, int ym) {
    if (2 * xm + 1 > 8 || 2 * ym + 1 > 8) return 0;
    const size_t bytes = (size_t)(RH_FT + 2 * xm) * (RH_FT + 2 * ym) * (8 * ns + RH_FPAD) * sizeof(float);
    return bytes <= 64 * 1024 ? bytes : 0;
}

extern "C" hipError_t rh_launch_box_masks(hipStream_t st, const FrameDev* fr, size_t nsamples,
                                          const uchar4* pixPattern, const float* pos2d, uint16_t* masks) {
    if (nsamples == 0) return hipSuccess;
    hipLaunchKernelGGL(rh_box_masks, dim3((unsigned)((nsamples + 255) / 256)), dim3(256), 0, st, fr, pixPattern,
                       pos2d, masks);
    return hipGetLastError();
}

// boxMask: the frame shape's masks from rh_launch_box_masks, or null when
// rh_filter_tile_lds is 0 for it (then the general kernel runs).
extern "C" hipError_t rh_launch_filter(hipStream_t st, const FrameDev* fr, int width, int height, int ns, int xm,
                                       int ym, const float4* samples, const uint16_t* boxMask,
                                       const uchar4* pixPattern, const float* pos2d, const float* filterTab,
                                       const float* dither, float* channels, uchar4* rgba8) {
    const int npix = width * height;
    if (npix <= 0) return hipSuccess;
    const size_t lds = boxMask ? rh_filter_tile_lds(ns, xm, ym) : 0;
    if (lds) {
        const int tilesX = (width + RH_FT - 1) / RH_FT, tilesY = (height + RH_FT - 1) / RH_FT;
        const int ntiles = tilesX * tilesY;
        const int per = (ntiles + 7) / 8;
        hipLaunchKernelGGL(rh_filter_tiled, dim3(8 * per), dim3(RH_FTHREADS), lds, st, fr, samples, boxMask, filterTab,
                           dither, channels, rgba8, tilesX, ntiles);
    } else {
        hipLaunchKernelGGL(rh_filter, dim3((npix + 255) / 256), dim3(256), 0, st, fr, samples, pixPattern, pos2d,
                           filterTab, dither, channels, rgba8);
    }
    return hipGetLastError();
}
