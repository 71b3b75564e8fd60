// siddon.h -- ray set-up and voxel traversal shared by the forward projector
// and its matched adjoint.
//
// Geometry follows CB_project_texture_chunked.cu:274-324 (aprict): positions
// are scaled by shrink_volume_factor, expressed in voxel units relative to the
// corner of voxel (0,0,0), and the integral runs along the finite segment from
// the source to the (sub-)pixel point.  The traversal itself is this library's
// own formulation:
//   * the ray is set up once in FP64 and re-centred on the middle of its chord
//     through the chunk, so every crossing parameter t is small;
//   * the crossing of plane P of axis a is the single FP32 expression
//        t_a(P) = fma(P - P0_a, r_a, c_a)
//     (P0_a an integer near the ray centre, r_a = 1/dir_a, c_a rounded once
//     from FP64), used identically for the chunk entry/exit and for every
//     step, so the forward projector and the adjoint produce bit-identical
//     weights and the traversal provably stays inside the chunk;
//   * one axis advances per step (ties resolved x, y, z), and the weight of a
//     voxel is the difference of consecutive crossing parameters.
#pragma once

#include "ap_common.h"

struct RayCfg {
    double tp[3];      // texpos (physical units)
    double svf;        // shrink_volume_factor
    double inv_vw;     // 1 / voxwid
    int lo[3], hi[3];  // chunk box in voxel indices
    int maxit;         // step bound: sum of chunk extents + 3
};

struct Ray {
    float r[3], c[3];  // t_a(P) = fma(P - P0_a, r_a, c_a)
    float pn[3];       // next plane of each axis, relative to P0_a
    float tn[3];       // parameter of that plane (AP_BIG if never crossed)
    float t_enter, t_exit;
    int idx[3];        // entry voxel (chunk-relative indices are idx - 0)
    int sg[3];         // +1 / -1 / 0 direction per axis
};

// physical point -> voxel units of the (whole) volume
__device__ __forceinline__ d3 to_vox(const RayCfg& rc, d3 p) {
    return d3{(p.x * rc.svf - rc.tp[0]) * rc.inv_vw,
              (p.y * rc.svf - rc.tp[1]) * rc.inv_vw,
              (p.z * rc.svf - rc.tp[2]) * rc.inv_vw};
}

// Pixel point of view g at fractional pixel indices (ju, kv) (pixel centre at
// integer values), CB_project_texture_chunked.cu:308-324.
__device__ __forceinline__ d3 pixel_point(const ViewGeom& g, double ju, double kv) {
    double a = ju - 0.5 * (double)(g.nu - 1);
    double b = kv - 0.5 * (double)(g.nv - 1);
    return d3{g.D.x + a * g.U.x + b * g.V.x,
              g.D.y + a * g.U.y + b * g.V.y,
              g.D.z + a * g.U.z + b * g.V.z};
}

__device__ __forceinline__ float tcross(float pn, float r, float c) {
    return __builtin_fmaf(pn, r, c);
}

// Sets up the ray S->Q (voxel units).  Returns false when the segment misses
// the chunk.
__device__ inline bool ray_setup(const RayCfg& rc, d3 S, d3 Q, Ray& R) {
    const double s[3] = {S.x, S.y, S.z};
    const double dd[3] = {Q.x - S.x, Q.y - S.y, Q.z - S.z};
    const double L = sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
    if (!(L > 0.0)) return false;
    const double inv_l = 1.0 / L;
    double u[3];
    double ta = 0.0, tb = L;
    for (int a = 0; a < 3; a++) {
        u[a] = dd[a] * inv_l;
        if (dd[a] == 0.0) {
            u[a] = 0.0;
            if (!(s[a] > rc.lo[a] && s[a] <= rc.hi[a])) return false;
        } else {
            double t1 = (rc.lo[a] - s[a]) / u[a], t2 = (rc.hi[a] - s[a]) / u[a];
            ta = fmax(ta, fmin(t1, t2));
            tb = fmin(tb, fmax(t1, t2));
        }
    }
    if (!(ta < tb)) return false;
    const double tm = 0.5 * (ta + tb);
    float te = (float)(-tm), tx = (float)(L - tm);
    double o[3], P0[3];
    for (int a = 0; a < 3; a++) {
        o[a] = s[a] + tm * u[a];
        if (u[a] == 0.0) {
            R.sg[a] = 0;
            R.r[a] = 0.f;
            R.c[a] = AP_BIG;
            R.pn[a] = 0.f;
            R.tn[a] = AP_BIG;
            R.idx[a] = (int)ceil(o[a]) - 1;   // dir==0 convention (lower voxel on a plane)
            P0[a] = 0.0;
            continue;
        }
        P0[a] = floor(o[a]);
        R.sg[a] = u[a] > 0.0 ? 1 : -1;
        R.r[a] = (float)(1.0 / u[a]);
        R.c[a] = (float)((P0[a] - o[a]) / u[a]);
        const int ein = u[a] > 0.0 ? rc.lo[a] : rc.hi[a];
        const int eout = u[a] > 0.0 ? rc.hi[a] : rc.lo[a];
        te = fmaxf(te, tcross((float)(ein - P0[a]), R.r[a], R.c[a]));
        tx = fminf(tx, tcross((float)(eout - P0[a]), R.r[a], R.c[a]));
    }
    if (!(te < tx)) return false;
    R.t_enter = te;
    R.t_exit = tx;
    for (int a = 0; a < 3; a++) {
        if (R.sg[a] == 0) continue;
        int i = (int)floor(o[a] + (double)te * u[a]);
        i = min(max(i, rc.lo[a]), rc.hi[a] - 1);
        int P = R.sg[a] > 0 ? i + 1 : i;
        const int pexit = R.sg[a] > 0 ? rc.hi[a] : rc.lo[a];
        float t = tcross((float)(P - P0[a]), R.r[a], R.c[a]);
        // the FP64 floor and the FP32 crossings can disagree by one plane
        // right at the entry point: step past planes that lie behind it
        for (int fix = 0; fix < 2 && t <= te && P != pexit; fix++) {
            i += R.sg[a];
            P += R.sg[a];
            t = tcross((float)(P - P0[a]), R.r[a], R.c[a]);
        }
        R.idx[a] = i;
        R.pn[a] = (float)(P - P0[a]);
        R.tn[a] = t;
    }
    return true;
}

// Line integral (voxel-length units) of the volume along a set-up ray.
template <typename Tv, typename IdxT>
__device__ __forceinline__ float ray_integral(const Tv* __restrict__ vol, const Ray& R,
                                              IdxT sx, IdxT sy, IdxT sz, int maxit) {
    IdxT off = (IdxT)R.idx[0] * sx + (IdxT)R.idx[1] * sy + (IdxT)R.idx[2] * sz;
    const IdxT dox = R.sg[0] >= 0 ? sx : -sx;
    const IdxT doy = R.sg[1] >= 0 ? sy : -sy;
    const IdxT doz = R.sg[2] >= 0 ? sz : -sz;
    const float gx = (float)R.sg[0], gy = (float)R.sg[1], gz = (float)R.sg[2];
    float tx = R.tn[0], ty = R.tn[1], tz = R.tn[2];
    float px = R.pn[0], py = R.pn[1], pz = R.pn[2];
    const float texit = R.t_exit;
    float tprev = R.t_enter, acc = 0.f;
    for (int it = 0; it < maxit; ++it) {
        const float tmn = fminf(fminf(tx, ty), tz);
        const float v = to_f32(vol[off]);
        acc = __builtin_fmaf(fminf(tmn, texit) - tprev, v, acc);
        if (tmn >= texit) break;
        tprev = tmn;
        if (tx == tmn) {
            px += gx;
            tx = tcross(px, R.r[0], R.c[0]);
            off += dox;
        } else if (ty == tmn) {
            py += gy;
            ty = tcross(py, R.r[1], R.c[1]);
            off += doy;
        } else {
            pz += gz;
            tz = tcross(pz, R.r[2], R.c[2]);
            off += doz;
        }
    }
    return acc;
}

// Visits the (voxel offset, weight) pairs of a set-up ray: the exact
// transpose of ray_integral (same steps, same weights).
template <typename IdxT, typename F>
__device__ __forceinline__ void ray_visit(const Ray& R, IdxT sx, IdxT sy, IdxT sz, int maxit, F&& f) {
    IdxT off = (IdxT)R.idx[0] * sx + (IdxT)R.idx[1] * sy + (IdxT)R.idx[2] * sz;
    const IdxT dox = R.sg[0] >= 0 ? sx : -sx;
    const IdxT doy = R.sg[1] >= 0 ? sy : -sy;
    const IdxT doz = R.sg[2] >= 0 ? sz : -sz;
    const float gx = (float)R.sg[0], gy = (float)R.sg[1], gz = (float)R.sg[2];
    float tx = R.tn[0], ty = R.tn[1], tz = R.tn[2];
    float px = R.pn[0], py = R.pn[1], pz = R.pn[2];
    const float texit = R.t_exit;
    float tprev = R.t_enter;
    for (int it = 0; it < maxit; ++it) {
        const float tmn = fminf(fminf(tx, ty), tz);
        f(off, fminf(tmn, texit) - tprev);
        if (tmn >= texit) break;
        tprev = tmn;
        if (tx == tmn) {
            px += gx;
            tx = tcross(px, R.r[0], R.c[0]);
            off += dox;
        } else if (ty == tmn) {
            py += gy;
            ty = tcross(py, R.r[1], R.c[1]);
            off += doy;
        } else {
            pz += gz;
            tz = tcross(pz, R.r[2], R.c[2]);
            off += doz;
        }
    }
}

int32_t make_ray_cfg(const ap_config_cb_project* cfg, const ap_vol_tex_chunk_spec* chunk,
                     const ap_array3* vol_desc, RayCfg& rc);
