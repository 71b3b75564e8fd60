// column.hip -- column-coherent Siddon forward projector and matched adjoint
// for views whose detector v axis is parallel to z (see column.h).
//
// Forward: a workgroup takes one detector (sub-)column of one view and 256
// consecutive rows.  It merges the column's x- and y-plane crossings into a
// segment list in LDS (each crossing computes its own rank, so the list is
// built by all threads at once), then every lane walks the list, inserting
// its own z crossings: per voxel a lane does one LDS broadcast read, a few
// FP32 operations and one gather.  Row-adjacent lanes read z-adjacent voxels
// of the same column, so each gather touches one or two cache lines.
//
// Adjoint: a workgroup owns a brick of kTX x kTY x kTZ voxels accumulated in
// LDS over all views.  Per view, each wave takes units (detector column,
// 64 rows) of the brick's footprint, builds that column's segment list inside
// the brick, and every lane scatters weight * y along its ray with LDS float
// atomics.  Segment boundaries and weights are those of the forward walk.
#include "column.h"

namespace {

constexpr int kXcdC = 8;
constexpr int kViewGroupC = 8;
constexpr int kRowsF = 256;        // rows per forward workgroup (4 waves)

struct SinoCfgC {
    uint32_t V, nu, nv;
    int64_t s0, s1, s2;
};

struct ColMap {
    uint32_t nkb;      // row blocks per column
    uint32_t nu;       // columns
    uint32_t total;    // padded tiles
    uint32_t per_xcd;
};

__device__ __forceinline__ float sub_off(uint32_t s, uint32_t ss) {
    return -0.5f + ((float)s + 0.5f) / (float)ss;
}

template <typename Tv, typename Tp, typename IdxT>
__global__ __launch_bounds__(kRowsF) void k_project_columns(
    RayCfg rc, SinoCfgC sc, ColMap cm, float mba, uint32_t ssu, uint32_t ssv,
    const ap_proj_spec* __restrict__ specs, const Tv* __restrict__ vol, IdxT vsx, IdxT vsy, IdxT vsz,
    Tp* __restrict__ sino, const Tp* __restrict__ b, uint32_t cap) {
    extern __shared__ __align__(16) char smem[];
    float* seg_h = reinterpret_cast<float*>(smem);
    IdxT* seg_o = reinterpret_cast<IdxT*>(smem + ((cap * sizeof(float) + 15) & ~size_t(15)));
    const uint32_t L = blockIdx.x;
    const uint32_t logical = (L % kXcdC) * cm.per_xcd + L / kXcdC;
    if (logical >= cm.total) return;
    const uint32_t kb = logical % cm.nkb;
    uint32_t r = logical / cm.nkb;
    const uint32_t vi = r % kViewGroupC;
    r /= kViewGroupC;
    const uint32_t j = r % cm.nu;
    const uint32_t p = (r / cm.nu) * kViewGroupC + vi;
    if (p >= sc.V || !view_is_upright(specs, p)) return;
    const int tid = threadIdx.x;
    const uint32_t k = kb * kRowsF + tid;
    const bool in = k < sc.nv;
    const ViewRays vr = make_view_rays(rc, specs, p);
    const int lo2[2] = {rc.lo[0], rc.lo[1]}, hi2[2] = {rc.hi[0], rc.hi[1]};
    float tot = 0.f;
    for (uint32_t su = 0; su < ssu; su++) {
        Col2D C;
        bool ok = col_setup(vr, rc.lo, rc.hi, (float)j + sub_off(su, ssu), C);
        float h0 = 0.f, h1 = 0.f;
        if (ok) {
            col_box(C, lo2, hi2, h0, h1);
            ok = h0 < h1;
        }
        if (!ok) continue;   // block-uniform
        const int ix0 = axis_index_at(C.r[0], C.c[0], C.P0[0], C.u[0], C.sg[0], C.ifix[0], h0, lo2[0], hi2[0]);
        const int iy0 = axis_index_at(C.r[1], C.c[1], C.P0[1], C.u[1], C.sg[1], C.ifix[1], h0, lo2[1], hi2[1]);
        const AxisSeq qx = axis_seq(C, 0, ix0, h1, lo2[0], hi2[0]);
        const AxisSeq qy = axis_seq(C, 1, iy0, h1, lo2[1], hi2[1]);
        const int nseg = qx.n + qy.n + 1;
        const float aux = fabsf(C.u[0]), auy = fabsf(C.u[1]);
        __syncthreads();   // the previous sub-column's list is consumed
        for (int m = tid; m < qx.n; m += kRowsF) {
            const float h = seq_at(qx, m);
            const int cy = seq_count(qy, h, true, auy);
            const int R = m + cy + 1;
            const int ix = min(max(ix0 + qx.sg * (m + 1), lo2[0]), hi2[0] - 1);
            const int iy = min(max(iy0 + qy.sg * cy, lo2[1]), hi2[1] - 1);
            seg_h[R] = h;
            seg_o[R] = (IdxT)ix * vsx + (IdxT)iy * vsy;
        }
        for (int l = tid; l < qy.n; l += kRowsF) {
            const float h = seq_at(qy, l);
            const int cx = seq_count(qx, h, false, aux);
            const int R = l + cx + 1;
            const int ix = min(max(ix0 + qx.sg * cx, lo2[0]), hi2[0] - 1);
            const int iy = min(max(iy0 + qy.sg * (l + 1), lo2[1]), hi2[1] - 1);
            seg_h[R] = h;
            seg_o[R] = (IdxT)ix * vsx + (IdxT)iy * vsy;
        }
        if (tid == 0) {
            seg_h[0] = h0;
            seg_o[0] = (IdxT)ix0 * vsx + (IdxT)iy0 * vsy;
            seg_h[nseg] = h1;
        }
        __syncthreads();
        if (!in) continue;
        float accu = 0.f;
        for (uint32_t sv = 0; sv < ssv; sv++) {
            LaneZ Z;
            if (!lane_setup(C, (float)k + sub_off(sv, ssv), rc.lo[2], rc.hi[2], Z)) continue;
            float zin, zout;
            lane_box(Z, rc.lo[2], rc.hi[2], zin, zout);
            const float lin = fmaxf(h0, zin), lout = fminf(h1, zout);
            if (!(lin < lout)) continue;
            const int iz = axis_index_at(Z.r, Z.c, Z.P0, Z.u, Z.sg, Z.ifix, lin, rc.lo[2], rc.hi[2]);
            float pz = Z.sg > 0 ? (float)(iz + 1) - Z.P0 : (float)iz - Z.P0;
            float hz = Z.sg == 0 ? AP_BIG : tcross(pz, Z.r, Z.c);
            const float gz = (float)Z.sg;
            IdxT zoff = (IdxT)iz * vsz;
            const IdxT dz = Z.sg >= 0 ? vsz : -vsz;
            float acc = 0.f;
            float hprev = h0;
            for (int s = 0; s < nseg; s++) {
                const float hn = seg_h[s + 1];
                const IdxT o = seg_o[s];
                float a = fmaxf(hprev, lin);
                const float bnd = fminf(hn, lout);
                while (hz < bnd) {
                    acc = __builtin_fmaf(fmaxf(hz - a, 0.f), to_f32(vol[o + zoff]), acc);
                    a = fmaxf(a, hz);
                    pz += gz;
                    hz = tcross(pz, Z.r, Z.c);
                    zoff += dz;
                }
                acc = __builtin_fmaf(fmaxf(bnd - a, 0.f), to_f32(vol[o + zoff]), acc);
                hprev = hn;
            }
            accu += acc * Z.sec;
        }
        tot += ssv > 1 ? accu / (float)ssv : accu;
    }
    if (!in) return;
    if (ssu > 1) tot /= (float)ssu;
    tot *= (float)(1.0 / rc.inv_vw);
    const int64_t o = (int64_t)p * sc.s0 + (int64_t)j * sc.s1 + (int64_t)k * sc.s2;
    const float base = to_f32(b ? b[o] : sino[o]);
    sino[o] = from_f32<Tp>(__builtin_fmaf(mba, tot, base));
}

// ------------------------------------------------------------------ adjoint
constexpr int kTX = 16, kTY = 16, kTZ = 64;
constexpr int kAdjT = 256;
constexpr int kWaves = kAdjT / 64;
constexpr int kVB = 16;            // views prepared per batch
constexpr int kSegCap = kTX + kTY + 2;

struct BrickViewC {
    ViewRays vr;
    float j0f, j1f;                // column range (fractional, before clamping)
    int j0, j1, k0, k1;
};

// pixel coordinates of the central projection (float), see project.hip
__device__ __forceinline__ bool proj_pt(const ViewRays& vd, const float X[3], float& ju, float& kv) {
    float S[3], Q[3], U[3], V[3];
    for (int a = 0; a < 3; a++) {
        S[a] = (float)vd.S[a];
        Q[a] = (float)vd.Q00[a];
        U[a] = (float)vd.U[a];
        V[a] = (float)vd.V[a];
    }
    const float n[3] = {U[1] * V[2] - U[2] * V[1], U[2] * V[0] - U[0] * V[2], U[0] * V[1] - U[1] * V[0]};
    float xs[3], qs[3];
    for (int a = 0; a < 3; a++) {
        xs[a] = X[a] - S[a];
        qs[a] = Q[a] - S[a];
    }
    const float den = xs[0] * n[0] + xs[1] * n[1] + xs[2] * n[2];
    const float num = qs[0] * n[0] + qs[1] * n[1] + qs[2] * n[2];
    if (!(den * num > 0.f)) return false;
    const float lam = num / den;
    float P[3];
    for (int a = 0; a < 3; a++) P[a] = S[a] + lam * xs[a] - Q[a];
    const float nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    const float pv[3] = {P[1] * V[2] - P[2] * V[1], P[2] * V[0] - P[0] * V[2], P[0] * V[1] - P[1] * V[0]};
    const float up[3] = {U[1] * P[2] - U[2] * P[1], U[2] * P[0] - U[0] * P[2], U[0] * P[1] - U[1] * P[0]};
    ju = (pv[0] * n[0] + pv[1] * n[1] + pv[2] * n[2]) / nn;
    kv = (up[0] * n[0] + up[1] * n[1] + up[2] * n[2]) / nn;
    return true;
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename Tv, typename Tp, typename Tm, typename IdxT>
__global__ __launch_bounds__(kAdjT) void k_backproject_columns(
    RayCfg rc, SinoCfgC sc, float scale, uint32_t ssu, uint32_t ssv,
    const ap_proj_spec* __restrict__ specs, Tv* __restrict__ vol, IdxT vsx, IdxT vsy, IdxT vsz,
    const Tm* __restrict__ mult, IdxT msx, IdxT msy, IdxT msz, const Tp* __restrict__ sino,
    uint32_t nbx, uint32_t nby, uint32_t nbz, uint32_t per_xcd) {
    __shared__ float acc[kTX * kTY * kTZ];
    __shared__ BrickViewC bv[kVB];
    __shared__ float seg_h[kWaves][kSegCap];
    __shared__ int seg_o[kWaves][kSegCap];
    const uint32_t L = blockIdx.x;
    const uint32_t logical = (L % kXcdC) * per_xcd + L / kXcdC;
    if (logical >= nbx * nby * nbz) return;
    const uint32_t bk = logical % nbz, bj = (logical / nbz) % nby, bi = logical / (nbz * nby);
    const int blo[3] = {rc.lo[0] + (int)bi * kTX, rc.lo[1] + (int)bj * kTY, rc.lo[2] + (int)bk * kTZ};
    const int bhi[3] = {min(blo[0] + kTX, rc.hi[0]), min(blo[1] + kTY, rc.hi[1]), min(blo[2] + kTZ, rc.hi[2])};
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int e = tid; e < kTX * kTY * kTZ; e += kAdjT) acc[e] = 0.f;
    const float cen[3] = {0.5f * (float)(rc.lo[0] + rc.hi[0]), 0.5f * (float)(rc.lo[1] + rc.hi[1]),
                          0.5f * (float)(rc.lo[2] + rc.hi[2])};
    float* my_h = seg_h[wave];
    int* my_o = seg_o[wave];
    for (uint32_t p0 = 0; p0 < sc.V; p0 += kVB) {
        const uint32_t nb = min((uint32_t)kVB, sc.V - p0);
        __syncthreads();
        if (tid < (int)nb) {
            BrickViewC& B = bv[tid];
            const uint32_t p = p0 + tid;
            B.j0 = 0; B.j1 = -1; B.k0 = 0; B.k1 = -1;
            if (view_is_upright(specs, p)) {
                B.vr = make_view_rays(rc, specs, p);
                float jmin = AP_BIG, jmax = -AP_BIG, kmin = AP_BIG, kmax = -AP_BIG;
                bool okp = true;
                for (int q = 0; q < 8; q++) {
                    const float X[3] = {(float)((q & 4) ? bhi[0] : blo[0]) - cen[0],
                                        (float)((q & 2) ? bhi[1] : blo[1]) - cen[1],
                                        (float)((q & 1) ? bhi[2] : blo[2]) - cen[2]};
                    float ju, kv;
                    if (!proj_pt(B.vr, X, ju, kv)) { okp = false; break; }
                    jmin = fminf(jmin, ju); jmax = fmaxf(jmax, ju);
                    kmin = fminf(kmin, kv); kmax = fmaxf(kmax, kv);
                }
                if (okp) {
                    B.j0 = max(0, (int)floorf(jmin - 0.501f));
                    B.j1 = min((int)sc.nu - 1, (int)ceilf(jmax + 0.501f));
                    B.k0 = max(0, (int)floorf(kmin - 0.501f));
                    B.k1 = min((int)sc.nv - 1, (int)ceilf(kmax + 0.501f));
                } else {
                    B.j0 = 0; B.j1 = (int)sc.nu - 1; B.k0 = 0; B.k1 = (int)sc.nv - 1;
                }
            }
        }
        __syncthreads();
        for (uint32_t q = 0; q < nb; q++) {
            const BrickViewC& B = bv[q];
            if (B.j1 < B.j0 || B.k1 < B.k0) continue;
            const uint32_t p = p0 + q;
            const int ncol = (B.j1 - B.j0 + 1) * (int)ssu;
            const int nchunk = (B.k1 - B.k0 + 64) / 64;
            const int nunits = ncol * nchunk;
            for (int un = wave; un < nunits; un += kWaves) {
                const int colx = un / nchunk, chunk = un % nchunk;
                const int j = B.j0 + colx / (int)ssu;
                const uint32_t su = (uint32_t)(colx % (int)ssu);
                Col2D C;
                if (!col_setup(B.vr, rc.lo, rc.hi, (float)j + sub_off(su, ssu), C)) continue;
                float h0, h1;
                col_box(C, blo, bhi, h0, h1);
                // the ray segment must also lie inside the chunk (it does: brick within chunk)
                if (!(h0 < h1)) continue;
                const int ix0 = axis_index_at(C.r[0], C.c[0], C.P0[0], C.u[0], C.sg[0], C.ifix[0], h0, blo[0], bhi[0]);
                const int iy0 = axis_index_at(C.r[1], C.c[1], C.P0[1], C.u[1], C.sg[1], C.ifix[1], h0, blo[1], bhi[1]);
                const AxisSeq qx = axis_seq(C, 0, ix0, h1, blo[0], bhi[0]);
                const AxisSeq qy = axis_seq(C, 1, iy0, h1, blo[1], bhi[1]);
                const int nseg = qx.n + qy.n + 1;
                wave_sync_lds();   // previous unit's list fully read
                if (lane < qx.n) {
                    const float h = seq_at(qx, lane);
                    const int cy = seq_count(qy, h, true, fabsf(C.u[1]));
                    const int ix = min(max(ix0 + qx.sg * (lane + 1), blo[0]), bhi[0] - 1);
                    const int iy = min(max(iy0 + qy.sg * cy, blo[1]), bhi[1] - 1);
                    my_h[lane + cy + 1] = h;
                    my_o[lane + cy + 1] = ((ix - blo[0]) * kTY + (iy - blo[1])) * kTZ;
                }
                if (lane < qy.n) {
                    const float h = seq_at(qy, lane);
                    const int cx = seq_count(qx, h, false, fabsf(C.u[0]));
                    const int ix = min(max(ix0 + qx.sg * cx, blo[0]), bhi[0] - 1);
                    const int iy = min(max(iy0 + qy.sg * (lane + 1), blo[1]), bhi[1] - 1);
                    my_h[lane + cx + 1] = h;
                    my_o[lane + cx + 1] = ((ix - blo[0]) * kTY + (iy - blo[1])) * kTZ;
                }
                if (lane == 0) {
                    my_h[0] = h0;
                    my_o[0] = ((ix0 - blo[0]) * kTY + (iy0 - blo[1])) * kTZ;
                    my_h[nseg] = h1;
                }
                wave_sync_lds();
                const int k = B.k0 + chunk * 64 + lane;
                if (k > B.k1) continue;
                const float y = to_f32(sino[(int64_t)p * sc.s0 + (int64_t)j * sc.s1 + (int64_t)k * sc.s2]);
                if (y == 0.f) continue;
                for (uint32_t sv = 0; sv < ssv; sv++) {
                    LaneZ Z;
                    if (!lane_setup(C, (float)k + sub_off(sv, ssv), rc.lo[2], rc.hi[2], Z)) continue;
                    float zin, zout;
                    lane_box(Z, blo[2], bhi[2], zin, zout);
                    const float lin = fmaxf(h0, zin), lout = fminf(h1, zout);
                    if (!(lin < lout)) continue;
                    const int iz = axis_index_at(Z.r, Z.c, Z.P0, Z.u, Z.sg, Z.ifix, lin, blo[2], bhi[2]);
                    float pz = Z.sg > 0 ? (float)(iz + 1) - Z.P0 : (float)iz - Z.P0;
                    float hz = Z.sg == 0 ? AP_BIG : tcross(pz, Z.r, Z.c);
                    const float gz = (float)Z.sg;
                    int zl = iz - blo[2];
                    const float wy = y * scale * Z.sec;
                    float hprev = h0;
                    for (int s = 0; s < nseg; s++) {
                        const float hn = my_h[s + 1];
                        const int o = my_o[s];
                        float a = fmaxf(hprev, lin);
                        const float bnd = fminf(hn, lout);
                        while (hz < bnd) {
                            const float w = fmaxf(hz - a, 0.f);
                            if (w > 0.f) atomicAdd(&acc[o + zl], w * wy);
                            a = fmaxf(a, hz);
                            pz += gz;
                            hz = tcross(pz, Z.r, Z.c);
                            zl += Z.sg;
                        }
                        const float w = fmaxf(bnd - a, 0.f);
                        if (w > 0.f) atomicAdd(&acc[o + zl], w * wy);
                        hprev = hn;
                    }
                }
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < kTX * kTY * kTZ; e += kAdjT) {
        const int li = e / (kTY * kTZ), lj = (e / kTZ) % kTY, lk = e % kTZ;
        const int gi = blo[0] + li, gj = blo[1] + lj, gk = blo[2] + lk;
        if (gi >= bhi[0] || gj >= bhi[1] || gk >= bhi[2]) continue;
        const float a = acc[e];
        if (a == 0.f) continue;
        const IdxT ov = (IdxT)gi * vsx + (IdxT)gj * vsy + (IdxT)gk * vsz;
        const float m = mult ? to_f32(mult[(IdxT)gi * msx + (IdxT)gj * msy + (IdxT)gk * msz]) : 1.f;
        vol[ov] = from_f32<Tv>(__builtin_fmaf(m, a, to_f32(vol[ov])));
    }
}

}  // namespace

// Launches the column-coherent forward projector (views with upright
// detectors; other views are left to the general kernel).  Returns
// AP_ERR_UNSUPPORTED when the x-y segment list would not fit in LDS.
int32_t launch_project_columns(const RayCfg& rc, uint32_t V, uint32_t nu, uint32_t nv, int64_t s0, int64_t s1,
                               int64_t s2, float mba, uint32_t ssu, uint32_t ssv, const ap_proj_spec* specs,
                               const ap_array3* vol_desc, const void* vol, int32_t sino_dtype, void* sino,
                               const void* b, bool idx64, hipStream_t s) {
    const uint32_t cap = (uint32_t)(rc.hi[0] - rc.lo[0]) + (uint32_t)(rc.hi[1] - rc.lo[1]) + 2;
    const size_t shm = ((cap * sizeof(float) + 15) & ~size_t(15)) + cap * (idx64 ? 8 : 4);
    if (shm > 60 * 1024) return AP_ERR_UNSUPPORTED;
    SinoCfgC sc{V, nu, nv, s0, s1, s2};
    ColMap cm;
    cm.nkb = (nv + kRowsF - 1) / kRowsF;
    cm.nu = nu;
    const uint64_t total = (uint64_t)((V + kViewGroupC - 1) / kViewGroupC) * kViewGroupC * nu * cm.nkb;
    if (total >= (1ull << 31)) return AP_ERR_UNSUPPORTED;
    cm.total = (uint32_t)total;
    cm.per_xcd = (cm.total + kXcdC - 1) / kXcdC;
    dim3 grid(cm.per_xcd * kXcdC);
    return dispatch_dtype(vol_desc->dtype, [&](auto tv) -> int32_t {
        using Tv = decltype(tv);
        if constexpr (std::is_same<Tv, double>::value) {
            return ap_fail(AP_ERR_UNSUPPORTED, "float64 volumes are not supported");
        } else {
            return dispatch_dtype(sino_dtype, [&](auto tp) -> int32_t {
                using Tp = decltype(tp);
                if constexpr (std::is_same<Tp, double>::value || std::is_same<Tp, bf16_t>::value) {
                    return ap_fail(AP_ERR_UNSUPPORTED, "sinogram dtype must be F32 or F16");
                } else {
                    if (idx64)
                        hipLaunchKernelGGL((k_project_columns<Tv, Tp, int64_t>), grid, dim3(kRowsF), shm, s, rc,
                                           sc, cm, mba, ssu, ssv, specs, (const Tv*)vol, vol_desc->strides[0],
                                           vol_desc->strides[1], vol_desc->strides[2], (Tp*)sino, (const Tp*)b,
                                           cap);
                    else
                        hipLaunchKernelGGL((k_project_columns<Tv, Tp, int32_t>), grid, dim3(kRowsF), shm, s, rc,
                                           sc, cm, mba, ssu, ssv, specs, (const Tv*)vol,
                                           (int32_t)vol_desc->strides[0], (int32_t)vol_desc->strides[1],
                                           (int32_t)vol_desc->strides[2], (Tp*)sino, (const Tp*)b, cap);
                    return ap_check_launch("project_columns");
                }
            });
        }
    });
}

int32_t launch_backproject_columns(const RayCfg& rc, uint32_t V, uint32_t nu, uint32_t nv, int64_t s0,
                                   int64_t s1, int64_t s2, float scale, uint32_t ssu, uint32_t ssv,
                                   const ap_proj_spec* specs, const ap_array3* vol_desc, void* vol,
                                   int32_t mdt, const int64_t* ms, const void* mult, int32_t sino_dtype,
                                   const void* sino, bool idx64, hipStream_t s) {
    SinoCfgC sc{V, nu, nv, s0, s1, s2};
    const uint32_t nbx = (uint32_t)(rc.hi[0] - rc.lo[0] + kTX - 1) / kTX;
    const uint32_t nby = (uint32_t)(rc.hi[1] - rc.lo[1] + kTY - 1) / kTY;
    const uint32_t nbz = (uint32_t)(rc.hi[2] - rc.lo[2] + kTZ - 1) / kTZ;
    const uint64_t nbr = (uint64_t)nbx * nby * nbz;
    if (nbr >= (1ull << 31)) return AP_ERR_UNSUPPORTED;
    const uint32_t per_xcd = (uint32_t)((nbr + kXcdC - 1) / kXcdC);
    dim3 grid(per_xcd * kXcdC);
    return dispatch_dtype(vol_desc->dtype, [&](auto tv) -> int32_t {
        using Tv = decltype(tv);
        return dispatch_dtype(sino_dtype, [&](auto tp) -> int32_t {
            using Tp = decltype(tp);
            return dispatch_dtype(mdt, [&](auto tm) -> int32_t {
                using Tm = decltype(tm);
                if constexpr (std::is_same<Tv, double>::value || std::is_same<Tp, double>::value ||
                              std::is_same<Tp, bf16_t>::value || std::is_same<Tm, double>::value) {
                    return ap_fail(AP_ERR_UNSUPPORTED, "unsupported dtype combination");
                } else {
                    if (idx64)
                        hipLaunchKernelGGL((k_backproject_columns<Tv, Tp, Tm, int64_t>), grid, dim3(kAdjT), 0, s,
                                           rc, sc, scale, ssu, ssv, specs, (Tv*)vol, vol_desc->strides[0],
                                           vol_desc->strides[1], vol_desc->strides[2], (const Tm*)mult, ms[0],
                                           ms[1], ms[2], (const Tp*)sino, nbx, nby, nbz, per_xcd);
                    else
                        hipLaunchKernelGGL((k_backproject_columns<Tv, Tp, Tm, int32_t>), grid, dim3(kAdjT), 0, s,
                                           rc, sc, scale, ssu, ssv, specs, (Tv*)vol, (int32_t)vol_desc->strides[0],
                                           (int32_t)vol_desc->strides[1], (int32_t)vol_desc->strides[2],
                                           (const Tm*)mult, (int32_t)ms[0], (int32_t)ms[1], (int32_t)ms[2],
                                           (const Tp*)sino, nbx, nby, nbz, per_xcd);
                    return ap_check_launch("backproject_columns");
                }
            });
        });
    });
}
