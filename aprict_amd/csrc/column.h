// column.h -- column-coherent Siddon traversal for detectors whose v axis is
// parallel to z (circular and helical scans with upright panels).
//
// For such a view every ray through detector column ju has the same
// projection onto the x-y plane: rays of one column differ only in their
// slope along z.  Parametrising a ray by the horizontal distance h along that
// common 2-D line, the x- and y-plane crossings are shared by the whole
// column and only the z-plane crossings depend on the row kv.  The exact
// Siddon integral of a ray is
//     sum over voxels of (h_out - h_in) * sec * f,   sec = sqrt(1 + m^2),
// m the ray's z slope per unit h.  The shared x-y crossings are merged once
// per column into a segment list in LDS -- every crossing's rank in the merged
// order is computed independently, so the list is built in parallel -- and
// each lane walks the list adding its own z crossings.
//
// All crossing parameters are single FP32 fma expressions of FP32 constants
// rounded once from FP64, shared by the forward projector and the adjoint, so
// both see bit-identical segment boundaries and weights.
#pragma once

#include "siddon.h"

// Per-column constants of the row z walks (FP64).
struct ColZ {
    double hc;                 // closest-approach parameter of the 2-D line
    double m0, dm;             // slope dz/dh of row kv: m0 + kv dm
    double zc;                 // source z plus the chunk's z centre (index units); lane_setup
                               // takes the z range [lo, hi) col_setup was given
};

// The shared 2-D line of one detector (sub-)column.
struct Col2D {
    float r[2], c[2], P0[2];   // h_a(P) = fma(P - P0_a, r_a, c_a), a = x, y
    float u[2];                // unit direction in x-y
    int sg[2];                 // direction signs (0: parallel to the planes)
    int ifix[2];               // index along an axis with sg == 0
    float hs, hd;              // source and pixel positions along h
    ColZ z;                    // what the rows of the column need
};

// Per-row z walk.
struct LaneZ {
    float r, c, P0;            // h_z(P) = fma(P - P0, r, c)
    float u;                   // dz/dh as float (position estimates)
    int sg, ifix;
    float sec;                 // 3-D length per unit h
};

// 2-D set-up of column ju (FP64, rounded once).  lo/hi: the chunk box.
// Returns false when the column's rays cannot meet the chunk in x-y.
__device__ inline bool col_setup(const ViewRays& vr, const int lo[3], const int hi[3], float ju, Col2D& C) {
    const double qx = vr.Q00[0] + (double)ju * vr.U[0];
    const double qy = vr.Q00[1] + (double)ju * vr.U[1];
    const double dx = qx - vr.S[0], dy = qy - vr.S[1];
    const double l = sqrt(dx * dx + dy * dy);
    if (!(l > 0.0)) return false;
    const double il = 1.0 / l;
    const double u[2] = {dx * il, dy * il};
    const double hc = -(vr.S[0] * u[0] + vr.S[1] * u[1]);
    C.z.hc = hc;
    C.z.m0 = (vr.Q00[2] + (double)ju * vr.U[2] - vr.S[2]) * il;
    C.z.dm = vr.V[2] * il;
    C.z.zc = vr.S[2] + 0.5 * (double)(lo[2] + hi[2]);
    C.hs = (float)(-hc);
    C.hd = (float)(l - hc);
    for (int a = 0; a < 2; a++) {
        const double oab = vr.S[a] + hc * u[a] + 0.5 * (double)(lo[a] + hi[a]);
        C.u[a] = (float)u[a];
        if (fabs(u[a]) < 1e-20) {
            C.sg[a] = 0;
            C.r[a] = 0.f;
            C.c[a] = AP_BIG;
            C.P0[a] = 0.f;
            if (!(oab > (double)lo[a] && oab <= (double)hi[a])) return false;
            C.ifix[a] = (int)ceil(oab) - 1;
            continue;
        }
        const double P0 = floor(oab);
        const double ru = 1.0 / u[a];
        C.P0[a] = (float)P0;
        C.sg[a] = u[a] > 0.0 ? 1 : -1;
        C.r[a] = (float)ru;
        C.c[a] = (float)((P0 - oab) * ru);
        C.ifix[a] = 0;
    }
    return true;
}

// Parameters [h0, h1) of the column line inside the x-y box [blo, bhi),
// intersected with the source->pixel segment.
__device__ __forceinline__ void col_box(const Col2D& C, const int blo[2], const int bhi[2], float& h0,
                                        float& h1) {
    h0 = C.hs;
    h1 = C.hd;
    for (int a = 0; a < 2; a++) {
        if (C.sg[a] == 0) {
            if (C.ifix[a] < blo[a] || C.ifix[a] >= bhi[a]) h1 = -AP_BIG;
            continue;
        }
        const float ein = (float)(C.sg[a] > 0 ? blo[a] : bhi[a]) - C.P0[a];
        const float eout = (float)(C.sg[a] > 0 ? bhi[a] : blo[a]) - C.P0[a];
        h0 = fmaxf(h0, tcross(ein, C.r[a], C.c[a]));
        h1 = fminf(h1, tcross(eout, C.r[a], C.c[a]));
    }
}

// Index along one axis after every plane whose crossing parameter is <= h,
// clamped to [blo, bhi).  (sg == 0: the fixed index.)
__device__ __forceinline__ int axis_index_at(float r, float c, float P0, float u, int sg, int ifix, float h,
                                             int blo, int bhi) {
    if (sg == 0) return ifix;
    int i = (int)(P0 + floorf((h - c) * u));
    i = min(max(i, blo), bhi - 1);
    if (sg > 0) {
        if (i + 1 < bhi && tcross((float)(i + 1) - P0, r, c) <= h) i++;
        if (i > blo && tcross((float)i - P0, r, c) > h) i--;
    } else {
        if (i > blo && tcross((float)i - P0, r, c) <= h) i--;
        if (i + 1 < bhi && tcross((float)(i + 1) - P0, r, c) > h) i++;
    }
    return i;
}

// Crossings of one axis strictly inside (h0, h1) along the column line,
// starting from the voxel i0 occupied at h0: the m-th is at
// tcross(first + sg*m - P0).
struct AxisSeq {
    float r, c, P0;
    int first, sg, n;
};

__device__ __forceinline__ float seq_at(const AxisSeq& q, int m) {
    return tcross((float)(q.first + q.sg * m) - q.P0, q.r, q.c);
}

__device__ inline AxisSeq axis_seq(const Col2D& C, int a, int i0, float h1, int blo, int bhi) {
    AxisSeq q;
    q.r = C.r[a];
    q.c = C.c[a];
    q.P0 = C.P0[a];
    q.sg = C.sg[a];
    if (C.sg[a] == 0) {
        q.first = 0;
        q.n = 0;
        return q;
    }
    q.first = C.sg[a] > 0 ? i0 + 1 : i0;
    // planes strictly before the box's exit plane
    const int avail = C.sg[a] > 0 ? bhi - q.first : q.first - blo;
    if (avail <= 0) {
        q.n = 0;
        return q;
    }
    const float x0 = seq_at(q, 0);
    int n = (int)floorf((h1 - x0) * fabsf(C.u[a])) + 1;
    n = min(max(n, 0), avail);
    while (n < avail && seq_at(q, n) < h1) n++;
    while (n > 0 && !(seq_at(q, n - 1) < h1)) n--;
    q.n = n;
    return q;
}

// Number of crossings of q strictly below h (strict) or at most h (!strict).
__device__ __forceinline__ int seq_count(const AxisSeq& q, float h, bool strict, float absu) {
    if (q.n == 0) return 0;
    int k = (int)floorf((h - seq_at(q, 0)) * absu) + 1;
    k = min(max(k, 0), q.n);
    if (strict) {
        while (k < q.n && seq_at(q, k) < h) k++;
        while (k > 0 && !(seq_at(q, k - 1) < h)) k--;
    } else {
        while (k < q.n && seq_at(q, k) <= h) k++;
        while (k > 0 && !(seq_at(q, k - 1) <= h)) k--;
    }
    return k;
}

// z walk of row kv of a column.  The slope m and the z offset of the
// row's closest approach, zab, are formed in FP64 (zab carries the crossing
// positions: an error e in zab moves every crossing by e/m); the plane base
// P0 = floor(zab) and P0 - zab are rounded once to FP32 and the reciprocal
// slope is a single FP32 rcp, so a crossing is off by at most a few 1e-7 of
// the ray length.  Forward and adjoint share this routine, hence identical
// constants.  Returns false for a ray that cannot meet [zlo, zhi)
// (horizontal rays outside it).
__device__ inline bool lane_setup(const ColZ& C, float kv, int zlo, int zhi, LaneZ& Z) {
    const double m = fma((double)kv, C.dm, C.m0);
    const double zab = fma(C.hc, m, C.zc);
    const float mf = (float)m;
    Z.sec = __builtin_amdgcn_sqrtf(__builtin_fmaf(mf, mf, 1.f));
    Z.u = mf;
    if (fabsf(mf) < 1e-20f) {
        Z.sg = 0;
        Z.r = 0.f;
        Z.c = AP_BIG;
        Z.P0 = 0.f;
        if (!(zab > (double)zlo && zab <= (double)zhi)) return false;
        Z.ifix = (int)ceil(zab) - 1;
        return true;
    }
    const double P0 = floor(zab);
    const float rm = __builtin_amdgcn_rcpf(mf);
    Z.P0 = (float)P0;
    Z.sg = mf > 0.f ? 1 : -1;
    Z.r = rm;
    Z.c = (float)(P0 - zab) * rm;
    Z.ifix = 0;
    return true;
}

// h range of the row inside the z slab [blo, bhi)
__device__ __forceinline__ void lane_box(const LaneZ& Z, int blo, int bhi, float& h0, float& h1) {
    if (Z.sg == 0) {
        if (Z.ifix < blo || Z.ifix >= bhi) { h0 = AP_BIG; h1 = -AP_BIG; }
        else { h0 = -AP_BIG; h1 = AP_BIG; }
        return;
    }
    const float ein = (float)(Z.sg > 0 ? blo : bhi) - Z.P0;
    const float eout = (float)(Z.sg > 0 ? bhi : blo) - Z.P0;
    h0 = tcross(ein, Z.r, Z.c);
    h1 = tcross(eout, Z.r, Z.c);
}
