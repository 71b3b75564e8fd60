// gather.h -- the per-view test and record shared by the gather-formulated
// matched adjoint (gather.hip) and the general adjoint (project.hip), which
// takes exactly the views the gather kernel does not.
#pragma once

#include "siddon.h"

// per-view record (whole-volume voxel coordinates)
struct GView {
    int32_t ok;          // 1: the view is handled by the gather kernel
    int32_t flip;        // z order: t = sub-row (0) or NT - 1 - sub-row (1)
    float sz;            // source height
    float dz0, dzs;      // dz_t = dz0 + t * dzs (dzs > 0)
    float sx, sy;        // source x, y
    float ax, ay;        // detector point of sub-column 0 minus the source (x, y)
    float bx, by;        // step of the detector point per sub-column (x, y)
    float orient;        // sign of cross(b, a): cross(b, X - S) * orient > 0 in front
    int32_t t0;          // rows (z order) whose ray does not rise (dz <= 0); set by the gather pre-pass
    float szlo;          // sz's FP32 rounding residual: the source height is sz + szlo
    float pad[2];
};
static_assert(sizeof(GView) == 64, "GView layout");

__device__ __forceinline__ float sub_offset_g(uint32_t s, uint32_t ss) {
    return -0.5f + ((float)s + 0.5f) / (float)ss;
}

// Whether view p is handled here: upright detector (v parallel to z, u
// without z component) and the chunk's x-y box strictly in front of the
// source.  Fills the view record.  Evaluated identically by the general
// adjoint, which takes every view for which it is false.
__device__ inline bool gather_view(const RayCfg& rc, const ap_proj_spec* __restrict__ specs, uint32_t p,
                                   uint32_t ssu, uint32_t ssv, uint32_t NT, GView& gv) {
    const ap_proj_spec* s = specs + p;
    gv.ok = 0;
    if (!(s->dspec.v.x == 0.f && s->dspec.v.y == 0.f && s->dspec.v.z != 0.f && s->dspec.u.z == 0.f)) return false;
    const ViewRays vr = make_view_rays(rc, specs, p);
    const double cen[3] = {0.5 * (rc.lo[0] + rc.hi[0]), 0.5 * (rc.lo[1] + rc.hi[1]), 0.5 * (rc.lo[2] + rc.hi[2])};
    const double Sx = vr.S[0] + cen[0], Sy = vr.S[1] + cen[1], Sz = vr.S[2] + cen[2];
    const double c0 = 0.5 / ssu - 0.5;
    const double ax = vr.Q00[0] + c0 * vr.U[0] - vr.S[0], ay = vr.Q00[1] + c0 * vr.U[1] - vr.S[1];
    const double bx = vr.U[0] / ssu, by = vr.U[1] / ssu;
    const double cba = bx * ay - by * ax;
    const double bn = sqrt(bx * bx + by * by);
    if (!(bn > 0.0) || cba == 0.0) return false;
    const double orient = cba > 0.0 ? 1.0 : -1.0;
    // every corner of the chunk's x-y box at least 1e-2 voxel in front of the source
    for (int q = 0; q < 4; q++) {
        const double X = (q & 1) ? rc.hi[0] : rc.lo[0], Y = (q & 2) ? rc.hi[1] : rc.lo[1];
        const double den = (bx * (Y - Sy) - by * (X - Sx)) * orient / bn;
        if (!(den > 1e-2)) return false;
    }
    // rays in z order: dz of sub-row kv is Q00z + kv Vz - Sz
    const double vz = vr.V[2];
    const int flip = vz < 0.0 ? 1 : 0;
    const double cv = 0.5 / ssv - 0.5;
    const double kv_first = flip ? (double)(NT - 1) / ssv + cv : cv;
    gv.ok = 1;
    gv.flip = flip;
    gv.sz = (float)Sz;
    gv.szlo = (float)(Sz - (double)gv.sz);
    gv.dz0 = (float)(vr.Q00[2] + kv_first * vz - vr.S[2]);
    gv.dzs = (float)(fabs(vz) / ssv);
    gv.sx = (float)Sx;
    gv.sy = (float)Sy;
    gv.ax = (float)ax;
    gv.ay = (float)ay;
    gv.bx = (float)bx;
    gv.by = (float)by;
    gv.orient = (float)orient;
    gv.t0 = 0;
    return true;
}


// Bytes of workspace the gather adjoint needs for V views (0 for an empty
// problem); whether it supports nv x ssv sub-rows; and its launch (returns
// AP_ERR_UNSUPPORTED when it cannot take the call).  With `left`, the launch
// also counts the views it does not take into the workspace and returns that
// device counter there.
uint64_t gather_workspace_bytes(uint32_t V, uint32_t nu, uint32_t nv, uint32_t ssu, uint32_t ssv);
bool gather_supported(uint32_t nv, uint32_t ssv);
int32_t launch_gather_adjoint(const RayCfg& rc, uint32_t V, uint32_t nu, uint32_t nv, int64_t s0, int64_t s1,
                              int64_t s2, uint32_t ssu, uint32_t ssv, double scale, const ap_proj_spec* specs,
                              const ap_array3* vol_desc, void* vol, int32_t mdt, const int64_t* ms,
                              const void* mult, int32_t sino_dtype, const void* sino, void* workspace,
                              uint64_t ws_bytes, hipStream_t s, const uint32_t** left);
