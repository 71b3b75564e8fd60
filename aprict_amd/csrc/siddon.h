// siddon.h -- ray set-up and voxel traversal shared by the forward projector
// and its matched adjoint.
//
// Geometry follows CB_project_texture_chunked.cu:274-324 (aprict): positions
// are scaled by shrink_volume_factor, expressed in voxel units relative to the
// corner of voxel (0,0,0), and the integral runs along the finite segment from
// the source to the (sub-)pixel point.  The traversal is this library's own
// formulation:
//   * per ray, FP64 code (voxel units relative to the chunk centre)
//     re-centres the ray on its closest approach to the chunk centre and
//     rounds the per-axis crossing constants once to FP32; the crossing of
//     plane P of axis a is then the single FP32 expression
//        t_a(P) = fma(P - P0_a, r_a, c_a)   (P0_a integer, r_a = 1/dir_a),
//     used identically for the chunk entry/exit, every step, and the brick
//     clipping of the adjoint, so forward and adjoint produce bit-identical
//     weights and every traversal provably stays inside its box;
//   * one axis advances per step (ties resolved x, y, z); a voxel's weight is
//     the difference of consecutive crossing parameters.
#pragma once

#include "ap_common.h"

struct RayCfg {
    double tp[3];      // texpos (physical units)
    double svf;        // shrink_volume_factor
    double inv_vw;     // 1 / voxwid
    int lo[3], hi[3];  // chunk box in voxel indices
    int maxit;         // step bound: sum of chunk extents + 3
};

// Per-view constants, voxel units, origin at the chunk centre (FP64).
struct ViewRays {
    double S[3];        // source
    double Q00[3];      // centre of pixel (0,0)
    double U[3], V[3];  // steps between adjacent pixel centres
};

struct Ray {
    float r[3], c[3];  // t_a(P) = fma(P - P0_a, r_a, c_a)
    float u[3];        // unit direction
    float P0[3];       // integer base plane per axis (as float)
    float pn[3];       // next plane of each axis, relative to P0_a
    float tn[3];       // parameter of that plane (AP_BIG if never crossed)
    float t_enter, t_exit;
    int idx[3];        // entry voxel (whole-volume indices)
    int sg[3];         // +1 / -1 / 0 direction per axis
};

__device__ __forceinline__ float tcross(float pn, float r, float c) {
    return __builtin_fmaf(pn, r, c);
}

// FP64 -> FP32 per-view constants (CB_project_texture_chunked.cu:308-324 for
// the pixel centres: centre of pixel (j,k) = pos + (j-(nu-1)/2) u + (k-(nv-1)/2) v).
__device__ inline ViewRays make_view_rays(const RayCfg& rc, const ap_proj_spec* __restrict__ specs,
                                          uint32_t p) {
    const ViewGeom g = load_view(specs, p);
    const double s = rc.svf * rc.inv_vw;
    ViewRays vr;
    const double cen[3] = {0.5 * (rc.lo[0] + rc.hi[0]), 0.5 * (rc.lo[1] + rc.hi[1]),
                           0.5 * (rc.lo[2] + rc.hi[2])};
    const double S[3] = {g.S.x, g.S.y, g.S.z}, D[3] = {g.D.x, g.D.y, g.D.z};
    const double U[3] = {g.U.x, g.U.y, g.U.z}, V[3] = {g.V.x, g.V.y, g.V.z};
    const double a = -0.5 * (double)(g.nu - 1), b = -0.5 * (double)(g.nv - 1);
    for (int q = 0; q < 3; q++) {
        const double tpv = rc.tp[q] * rc.inv_vw;
        vr.S[q] = S[q] * s - tpv - cen[q];
        vr.Q00[q] = (D[q] + a * U[q] + b * V[q]) * s - tpv - cen[q];
        vr.U[q] = U[q] * s;
        vr.V[q] = V[q] * s;
    }
    return vr;
}

// Ray set-up for the sub-pixel point at fractional pixel indices (ju, kv):
// FP64 up to the per-axis crossing constants, which are rounded once to FP32.
// lo/hi: the chunk box (whole-volume voxel indices).  Returns false when the
// segment misses the chunk.
__device__ inline bool ray_setup(const ViewRays& vr, const int lo[3], const int hi[3], float ju,
                                 float kv, Ray& R) {
    double d[3];
    for (int a = 0; a < 3; a++) {
        const double q = vr.Q00[a] + (double)ju * vr.U[a] + (double)kv * vr.V[a];
        d[a] = q - vr.S[a];
    }
    const double l2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (!(l2 > 0.0)) return false;
    const double len = sqrt(l2);
    double u[3];
    for (int a = 0; a < 3; a++) u[a] = d[a] / len;
    // closest approach to the chunk centre: parameter tc (distance from the source)
    const double tc = -(vr.S[0] * u[0] + vr.S[1] * u[1] + vr.S[2] * u[2]);
    float te = (float)(-tc), tx = (float)(len - tc);
    for (int a = 0; a < 3; a++) {
        // closest point, whole-volume voxel coordinates
        const double oab = vr.S[a] + tc * u[a] + 0.5 * (double)(lo[a] + hi[a]);
        R.u[a] = (float)u[a];
        if (fabs(u[a]) < 1e-20) {
            R.sg[a] = 0;
            R.r[a] = 0.f;
            R.c[a] = AP_BIG;
            R.pn[a] = 0.f;
            R.tn[a] = AP_BIG;
            R.P0[a] = 0.f;
            if (!(oab > (double)lo[a] && oab <= (double)hi[a])) return false;
            R.idx[a] = (int)ceil(oab) - 1;   // lower voxel on a plane (reference convention)
            continue;
        }
        const double P0 = floor(oab);
        R.P0[a] = (float)P0;
        R.sg[a] = u[a] > 0.0 ? 1 : -1;
        R.r[a] = (float)(1.0 / u[a]);
        R.c[a] = (float)((P0 - oab) / u[a]);
        const float ein = (float)(R.sg[a] > 0 ? lo[a] : hi[a]) - R.P0[a];
        const float eout = (float)(R.sg[a] > 0 ? hi[a] : lo[a]) - R.P0[a];
        te = fmaxf(te, tcross(ein, R.r[a], R.c[a]));
        tx = fminf(tx, tcross(eout, R.r[a], R.c[a]));
    }
    if (!(te < tx)) return false;
    R.t_enter = te;
    R.t_exit = tx;
    return true;
}

// Places the traversal state at parameter t0 (>= t_enter) inside the box
// [blo, bhi): per axis, the voxel after every plane whose crossing parameter
// is <= t0 (the DDA state after processing all crossings up to t0), clamped
// to the box, and the next plane beyond it.
__device__ inline void ray_start(Ray& R, float t0, const int blo[3], const int bhi[3]) {
    for (int a = 0; a < 3; a++) {
        if (R.sg[a] == 0) continue;
        // voxel containing the point at t0, from its position relative to P0
        int i = (int)(R.P0[a] + floorf((t0 - R.c[a]) * R.u[a]));
        i = min(max(i, blo[a]), bhi[a] - 1);
        const float P0 = R.P0[a];
        if (R.sg[a] > 0) {
            // next plane i+1 must lie beyond t0, previous plane i at or before it
            if (i + 1 < bhi[a] && tcross((float)(i + 1) - P0, R.r[a], R.c[a]) <= t0) i++;
            if (i > blo[a] && tcross((float)i - P0, R.r[a], R.c[a]) > t0) i--;
            R.pn[a] = (float)(i + 1) - P0;
        } else {
            if (i > blo[a] && tcross((float)i - P0, R.r[a], R.c[a]) <= t0) i--;
            if (i + 1 < bhi[a] && tcross((float)(i + 1) - P0, R.r[a], R.c[a]) > t0) i++;
            R.pn[a] = (float)i - P0;
        }
        R.idx[a] = i;
        R.tn[a] = tcross(R.pn[a], R.r[a], R.c[a]);
    }
}

// Crossing parameters of the box [blo, bhi) along the ray: [tb0, tb1).
__device__ __forceinline__ void ray_box(const Ray& R, const int blo[3], const int bhi[3], float& tb0,
                                        float& tb1) {
    tb0 = R.t_enter;
    tb1 = R.t_exit;
    for (int a = 0; a < 3; a++) {
        if (R.sg[a] == 0) {
            if (R.idx[a] < blo[a] || R.idx[a] >= bhi[a]) tb1 = -AP_BIG;
            continue;
        }
        const float ein = (float)(R.sg[a] > 0 ? blo[a] : bhi[a]) - R.P0[a];
        const float eout = (float)(R.sg[a] > 0 ? bhi[a] : blo[a]) - R.P0[a];
        tb0 = fmaxf(tb0, tcross(ein, R.r[a], R.c[a]));
        tb1 = fminf(tb1, tcross(eout, R.r[a], R.c[a]));
    }
}

// One traversal step: weight of the current voxel, then advance (unless the
// segment ended).  Returns false once the end parameter is reached.
template <typename IdxT>
struct Walker {
    float tx, ty, tz, px, py, pz, tprev, tend;
    float rx, ry, rz, cx, cy, cz, gx, gy, gz;
    IdxT off, dox, doy, doz;

    __device__ __forceinline__ void init(const Ray& R, float t0, float t1, IdxT sx, IdxT sy, IdxT sz, IdxT base) {
        tx = R.tn[0]; ty = R.tn[1]; tz = R.tn[2];
        px = R.pn[0]; py = R.pn[1]; pz = R.pn[2];
        rx = R.r[0]; ry = R.r[1]; rz = R.r[2];
        cx = R.c[0]; cy = R.c[1]; cz = R.c[2];
        gx = (float)R.sg[0]; gy = (float)R.sg[1]; gz = (float)R.sg[2];
        off = base;
        dox = R.sg[0] >= 0 ? sx : -sx;
        doy = R.sg[1] >= 0 ? sy : -sy;
        doz = R.sg[2] >= 0 ? sz : -sz;
        tprev = t0;
        tend = t1;
    }

    // weight w of voxel `off`; advances state; false when this was the last voxel
    __device__ __forceinline__ bool step(float& w, IdxT& o) {
        const float tmn = fminf(fminf(tx, ty), tz);
        const float tcl = fminf(tmn, tend);
        w = fmaxf(tcl - tprev, 0.f);
        o = off;
        tprev = tcl;
        if (tmn >= tend) return false;
        if (tx == tmn) {
            px += gx; tx = tcross(px, rx, cx); off += dox;
        } else if (ty == tmn) {
            py += gy; ty = tcross(py, ry, cy); off += doy;
        } else {
            pz += gz; tz = tcross(pz, rz, cz); off += doz;
        }
        return true;
    }
};

// True when the view's pixel v vector is parallel to z (x = y = 0 exactly).
__device__ __forceinline__ bool view_is_upright(const ap_proj_spec* __restrict__ specs, uint32_t p) {
    const ap_proj_spec* s = specs + p;
    return s->dspec.v.x == 0.f && s->dspec.v.y == 0.f && s->dspec.v.z != 0.f;
}

int32_t make_ray_cfg(const ap_config_cb_project* cfg, const ap_vol_tex_chunk_spec* chunk,
                     const ap_array3* vol_desc, RayCfg& rc);
