// column.hip -- column-coherent Siddon forward projector for views whose
// detector v axis is parallel to z (see column.h).
//
// A workgroup takes one detector (sub-)column of one view and 128 consecutive
// rows (64 or 256 in the other launch-shape variants).  It merges the
// column's x- and y-plane crossings into a segment list in LDS (each crossing
// computes its own rank, so the list is built by all threads at once), then
// every lane walks the list, inserting its own z crossings: per voxel a lane
// does one LDS broadcast read, a few FP32 operations and one gather.
// Row-adjacent lanes read z-adjacent voxels of the same column, so each gather
// touches one or two cache lines.  (The matched adjoint of these views is the
// gather formulation in gather.hip.)
#include "column.h"

namespace {

constexpr int kXcdC = 8;
constexpr int kViewGroupC = 8;
// detector rows per forward workgroup (AP_TUNE_PROJECT_COLUMN variants);
// waves always hold 64 consecutive rows, so every variant gives identical sums
constexpr int kColRows[3] = {128, 256, 64};

struct SinoCfgC {
    uint32_t V, nu, nv;
    int64_t s0, s1, s2;
    int64_t b0, b1, b2;   // element strides of the fused residual's b (forward)
    uint32_t vec16;    // f32 or f16, v contiguous, rows 16-byte aligned, nv a multiple of 16 bytes
    uint32_t o32;      // every element offset fits 31 bits: 32-bit address arithmetic
};

// element offset of sinogram sample (p, j, k)
__device__ __forceinline__ int64_t sino_off(const SinoCfgC& sc, uint32_t p, int j, int k) {
    if (sc.o32)
        return (int64_t)(uint32_t)((int32_t)p * (int32_t)sc.s0 + j * (int32_t)sc.s1 + k * (int32_t)sc.s2);
    return (int64_t)p * sc.s0 + (int64_t)j * sc.s1 + (int64_t)k * sc.s2;
}

struct ColMap {
    uint32_t nkb;      // row blocks per column
    uint32_t nu;       // columns
    uint32_t total;    // padded tiles
    uint32_t per_xcd;
};

__device__ __forceinline__ float sub_off(uint32_t s, uint32_t ss) {
    return -0.5f + ((float)s + 0.5f) / (float)ss;
}

// Volume element at offset o: bytes for 32-bit offsets, elements for 64-bit.
// Volume element at 32-bit byte offset o, unconverted.
template <typename Tv>
__device__ __forceinline__ Tv vraw(const Tv* __restrict__ v, uint32_t o) {
    return *reinterpret_cast<const Tv*>(reinterpret_cast<const char*>(v) + o);
}

template <typename Tv, typename OffT>
__device__ __forceinline__ float vget(const Tv* __restrict__ v, OffT o) {
    if constexpr (sizeof(OffT) == 4)
        return to_f32(*reinterpret_cast<const Tv*>(reinterpret_cast<const char*>(v) + o));
    else
        return to_f32(v[o]);
}

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Minimum / maximum over lanes 0-7 of a wave (all lanes active), as a
// wave-uniform value: quad permutations [1,0,3,2] and [2,3,0,1] and a mirror
// within each half row of 8 lanes (DPP moves, no LDS), then lane 0's value.
template <int kCtrl>
__device__ __forceinline__ float dpp_mov(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), kCtrl, 0xF, 0xF, false));
}

__device__ __forceinline__ float lanes8_min(float v) {
    v = fminf(v, dpp_mov<0xB1>(v));
    v = fminf(v, dpp_mov<0x4E>(v));
    v = fminf(v, dpp_mov<0x141>(v));
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

__device__ __forceinline__ float lanes8_max(float v) {
    v = fmaxf(v, dpp_mov<0xB1>(v));
    v = fmaxf(v, dpp_mov<0x4E>(v));
    v = fmaxf(v, dpp_mov<0x141>(v));
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// First index s in [0, n) with he[s] > x (n if none); he is non-decreasing.
__device__ __forceinline__ int first_above(const float* he, int n, float x) {
    int a = 0, b = n;
    while (a < b) {
        const int m = (a + b) >> 1;
        if (he[m] > x) b = m;
        else a = m + 1;
    }
    return a;
}

// Quotient n / d for a wave-uniform divisor d >= 1 and 0 <= n < 2^22: a
// float reciprocal estimate and one integer correction.
struct UDiv {
    float inv;
    int d;
};

__device__ __forceinline__ UDiv make_udiv(int d) { return UDiv{1.f / (float)d, d}; }

__device__ __forceinline__ int udiv(int n, const UDiv& D) {
    int q = (int)((float)n * D.inv);
    const int r = n - q * D.d;
    q += (r >= D.d) - (r < 0);
    return q;
}

// One row's z walk along segments [s0, s1) of a column list (he: segment
// end parameters, ob: x-y offsets).  hp starts at the row's entry t0; every
// z crossing before a segment's clamped end splits that segment.  Weights
// are differences of the same FP32 values the adjoint uses.
template <typename Tv, typename OffT>
__device__ __forceinline__ float walk_rows_general(const Tv* __restrict__ vol, const float* he, const OffT* ob,
                                                   int s0, int s1, float t0, float t1, float hz, float pz,
                                                   float gz, float rz, float cz, OffT zoff, OffT dz) {
    float acc = 0.f, hp = t0;
    for (int s = s0; s < s1; s++) {
        const float ce = __builtin_amdgcn_fmed3f(he[s], t0, t1);
        const OffT o = ob[s];
        while (hz < ce) {
            acc = __builtin_fmaf(hz - hp, vget(vol, (OffT)(o + zoff)), acc);
            hp = hz;
            pz += gz;
            hz = tcross(pz, rz, cz);
            zoff += dz;
        }
        acc = __builtin_fmaf(ce - hp, vget(vol, (OffT)(o + zoff)), acc);
        hp = ce;
    }
    return acc;
}

// The same walk when no segment holds more than one z crossing of any row
// of the wave: straight-line code over 8 segments at a time, so that up to
// 16 gathers are in flight per lane.  The voxel after a z crossing is
// gathered only in lanes that crossed (others add a zero weight times 0):
// the masked gathers relieve the memory pipeline, about 4% of the kernel.
// The segment list is read with wave-uniform addresses (broadcast).
template <typename Tv>
__device__ __forceinline__ float walk_rows_single(const Tv* __restrict__ vol, const float* he, const uint32_t* ob,
                                                  int s0, int s1, float t0, float t1, float hz, float pz,
                                                  float gz, float rz, float cz, uint32_t zoff, uint32_t dz) {
    float acc = 0.f, hp = t0;
    for (int s = s0 & ~3; s < s1; s += 8) {
        const float4 h0 = *reinterpret_cast<const float4*>(he + s);
        const float4 h1 = *reinterpret_cast<const float4*>(he + s + 4);
        const uint4 o0 = *reinterpret_cast<const uint4*>(ob + s);
        const uint4 o1 = *reinterpret_cast<const uint4*>(ob + s + 4);
        const float hv[8] = {h0.x, h0.y, h0.z, h0.w, h1.x, h1.y, h1.z, h1.w};
        const uint32_t ov[8] = {o0.x, o0.y, o0.z, o0.w, o1.x, o1.y, o1.z, o1.w};
        float wb[8], wa[8];
        Tv vb[8], va[8];   // raw values, converted in the accumulation below
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const float ce = __builtin_amdgcn_fmed3f(hv[q], t0, t1);
            const bool cr = hz < ce;
            const float c = cr ? hz : ce;
            wb[q] = c - hp;
            wa[q] = ce - c;
            hp = ce;
            const uint32_t a0 = ov[q] + zoff;
            zoff += cr ? dz : 0u;
            pz += cr ? gz : 0.f;
            hz = tcross(pz, rz, cz);
            vb[q] = vraw(vol, a0);
            // the gather after a crossing only where there is one (Tv{} is
            // 0.0).  The zero is opaque: otherwise the compiler folds the
            // conversion of 16-bit values into the branch, whose wait for the
            // load then stalls every gather in flight
            Tv zero = Tv{};
            asm volatile("" : "+v"(zero));
            va[q] = cr ? vraw(vol, ov[q] + zoff) : zero;
        }
#pragma unroll
        for (int q = 0; q < 8; q++) {
            acc = __builtin_fmaf(wb[q], to_f32(vb[q]), acc);
            acc = __builtin_fmaf(wa[q], to_f32(va[q]), acc);
        }
    }
    return acc;
}

// Rows of a wave whose z slope |m| stays below this never cross two z planes
// inside one x-y segment: segments are at most sqrt(2) long in h and z planes
// are 1/|m| apart.
constexpr float kSingleSlope = 0.6f;

// Merged x-y crossing list of column C inside the box [lo2, hi2), built by
// the nthr threads of the block: seg_he[s] is the end parameter of segment s,
// seg_ob[s] the x-y offset of its voxel; 8 zero-length padding segments
// follow the last one.  Returns the number of segments.  The caller
// synchronises before (the previous list is consumed) and after.
template <typename OffT>
__device__ inline int build_column_list(const Col2D& C, float h0, float h1, const int lo2[2], const int hi2[2],
                                        OffT vsx, OffT vsy, float* seg_he, OffT* seg_ob, int tid, int nthr) {
    const int ix0 = axis_index_at(C.r[0], C.c[0], C.P0[0], C.u[0], C.sg[0], C.ifix[0], h0, lo2[0], hi2[0]);
    const int iy0 = axis_index_at(C.r[1], C.c[1], C.P0[1], C.u[1], C.sg[1], C.ifix[1], h0, lo2[1], hi2[1]);
    const AxisSeq qx = axis_seq(C, 0, ix0, h1, lo2[0], hi2[0]);
    const AxisSeq qy = axis_seq(C, 1, iy0, h1, lo2[1], hi2[1]);
    const int nseg = qx.n + qy.n + 1;
    const float aux = fabsf(C.u[0]), auy = fabsf(C.u[1]);
    // segment R-1 ends at the crossing of rank R; segment R starts there
    for (int m = tid; m < qx.n; m += nthr) {
        const float h = seq_at(qx, m);
        const int cy = seq_count(qy, h, true, auy);
        const int R = m + cy + 1;
        const int ix = min(max(ix0 + qx.sg * (m + 1), lo2[0]), hi2[0] - 1);
        const int iy = min(max(iy0 + qy.sg * cy, lo2[1]), hi2[1] - 1);
        seg_he[R - 1] = h;
        seg_ob[R] = (OffT)ix * vsx + (OffT)iy * vsy;
    }
    for (int l = tid; l < qy.n; l += nthr) {
        const float h = seq_at(qy, l);
        const int cx = seq_count(qx, h, false, aux);
        const int R = l + cx + 1;
        const int ix = min(max(ix0 + qx.sg * cx, lo2[0]), hi2[0] - 1);
        const int iy = min(max(iy0 + qy.sg * (l + 1), lo2[1]), hi2[1] - 1);
        seg_he[R - 1] = h;
        seg_ob[R] = (OffT)ix * vsx + (OffT)iy * vsy;
    }
    if (tid == 0) seg_ob[0] = (OffT)ix0 * vsx + (OffT)iy0 * vsy;
    if (tid < 9) {
        // the last segment ends at h1; the padding repeats the final voxel so
        // the unrolled walks may read up to 7 entries past the end
        const int ixe = min(max(ix0 + qx.sg * qx.n, lo2[0]), hi2[0] - 1);
        const int iye = min(max(iy0 + qy.sg * qy.n, lo2[1]), hi2[1] - 1);
        seg_he[nseg - 1 + tid] = h1;
        if (tid > 0) seg_ob[nseg - 1 + tid] = (OffT)ixe * vsx + (OffT)iye * vsy;
    }
    return nseg;
}

// z walk state of one row of a column: entry t0 / exit t1 of the chunk, the
// voxel index at t0 and the next z crossing.  Inactive rows get an empty
// range at h0 and an in-range index, so they may run the walks unmasked.
struct RowWalk {
    float t0, t1, hz, pz, gz, rz, cz, sec, slope;
    int iz, sg;
    bool act;
};

__device__ inline RowWalk row_setup(const Col2D& C, float h0, float h1, float kv, bool in, int zlo, int zhi) {
    RowWalk W;
    W.t0 = h0; W.t1 = h0; W.hz = AP_BIG; W.pz = 0.f; W.gz = 0.f; W.rz = 0.f; W.cz = AP_BIG;
    W.sec = 0.f; W.slope = 0.f; W.iz = zlo; W.sg = 0; W.act = false;
    LaneZ Z;
    if (!in || !lane_setup(C.z, kv, zlo, zhi, Z)) return W;
    float zin, zout;
    lane_box(Z, zlo, zhi, zin, zout);
    const float lin = fmaxf(h0, zin), lout = fminf(h1, zout);
    if (!(lin < lout)) return W;
    W.act = true;
    W.t0 = lin;
    W.t1 = lout;
    W.sec = Z.sec;
    W.slope = fabsf(Z.u);
    W.sg = Z.sg;
    W.iz = axis_index_at(Z.r, Z.c, Z.P0, Z.u, Z.sg, Z.ifix, lin, zlo, zhi);
    if (Z.sg != 0) {
        W.pz = Z.sg > 0 ? (float)(W.iz + 1) - Z.P0 : (float)W.iz - Z.P0;
        W.gz = (float)Z.sg;
        W.rz = Z.r;
        W.cz = Z.c;
        W.hz = tcross(W.pz, W.rz, W.cz);
    }
    return W;
}

// The walk's gathers and crossing chain are latency that only more waves
// hide.  kOne: one sample per pixel (ssu = ssv = 1), so nothing of the
// column's set-up stays live across the walk for a next sub-sample: 50 VGPRs,
// 8 waves per SIMD.  With sub-samples the kernel is held to 5 waves (96
// VGPRs; the set-up spills 14 registers to scratch once per column).
// 512^3 x 720 views, 128 rows: 4 waves 69.7 ms, 5 waves 65.4 ms, kOne 56.0 ms.
template <int kRowsF, typename Tv, typename Tp, typename OffT, bool kOne>
__global__ __launch_bounds__(kRowsF, 5) void k_project_columns(
    RayCfg rc, SinoCfgC sc, ColMap cm, float mba, uint32_t ssu, uint32_t ssv,
    const ap_proj_spec* __restrict__ specs, const Tv* __restrict__ vol, OffT vsx, OffT vsy, OffT vsz,
    Tp* __restrict__ sino, const Tp* __restrict__ b, uint32_t cap) {
    extern __shared__ __align__(16) char smem[];
    float* seg_he = reinterpret_cast<float*>(smem);
    OffT* seg_ob = reinterpret_cast<OffT*>(smem + ((cap * sizeof(float) + 15) & ~size_t(15)));
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
    if constexpr (kOne) ssu = ssv = 1;
    for (uint32_t su = 0; su < ssu; su++) {
        Col2D C;
        bool ok = col_setup(vr, rc.lo, rc.hi, (float)j + sub_off(su, ssu), C);
        float h0 = 0.f, h1 = 0.f;
        if (ok) {
            col_box(C, lo2, hi2, h0, h1);
            ok = h0 < h1;
        }
        if (!ok) continue;   // block-uniform
        __syncthreads();   // the previous sub-column's list is consumed
        const int nseg = build_column_list(C, h0, h1, lo2, hi2, vsx, vsy, seg_he, seg_ob, tid, kRowsF);
        __syncthreads();
        float accu = 0.f;
        for (uint32_t sv = 0; sv < ssv; sv++) {
            const RowWalk W = row_setup(C, h0, h1, (float)k + sub_off(sv, ssv), in, rc.lo[2], rc.hi[2]);
            const float wt0 = wave_min(W.act ? W.t0 : AP_BIG);
            const float wt1 = wave_max(W.act ? W.t1 : -AP_BIG);
            if (!(wt0 < wt1)) continue;   // wave-uniform
            const int s0 = first_above(seg_he, nseg, wt0);
            const int s1 = min(first_above(seg_he, nseg, wt1) + 1, nseg);
            const bool steep = __ballot(W.act && W.slope >= kSingleSlope) != 0;
            const OffT zoff = (OffT)W.iz * vsz;
            const OffT dz = W.sg >= 0 ? vsz : (OffT)(OffT(0) - vsz);
            float acc;
            if constexpr (sizeof(OffT) == 4) {
                if (!steep)
                    acc = walk_rows_single(vol, seg_he, seg_ob, s0, s1, W.t0, W.t1, W.hz, W.pz, W.gz, W.rz, W.cz,
                                           zoff, dz);
                else
                    acc = walk_rows_general(vol, seg_he, seg_ob, s0, s1, W.t0, W.t1, W.hz, W.pz, W.gz, W.rz, W.cz,
                                            zoff, dz);
            } else {
                acc = walk_rows_general(vol, seg_he, seg_ob, s0, s1, W.t0, W.t1, W.hz, W.pz, W.gz, W.rz, W.cz, zoff,
                                        dz);
            }
            if (W.act) accu += acc * W.sec;
        }
        tot += ssv > 1 ? accu / (float)ssv : accu;
    }
    if (!in) return;
    if (ssu > 1) tot /= (float)ssu;
    tot *= (float)(1.0 / rc.inv_vw);
    const int64_t o = (int64_t)p * sc.s0 + (int64_t)j * sc.s1 + (int64_t)k * sc.s2;
    const float base =
        to_f32(b ? b[(int64_t)p * sc.b0 + (int64_t)j * sc.b1 + (int64_t)k * sc.b2] : sino[o]);
    sino[o] = from_f32<Tp>(__builtin_fmaf(mba, tot, base));
}

}  // namespace

// Launches the column-coherent forward projector (views with upright
// detectors; other views are left to the general kernel).  Returns
// AP_ERR_UNSUPPORTED when the x-y segment list would not fit in LDS.
int32_t launch_project_columns(const RayCfg& rc, uint32_t V, uint32_t nu, uint32_t nv, int64_t s0, int64_t s1,
                               int64_t s2, float mba, uint32_t ssu, uint32_t ssv, const ap_proj_spec* specs,
                               const ap_array3* vol_desc, const void* vol, int32_t sino_dtype, void* sino,
                               const void* b, const int64_t* bs, bool idx64, hipStream_t s) {
    // segment list: at most Nx + Ny + 1 segments plus 8 padding entries
    const uint32_t cap = ((uint32_t)(rc.hi[0] - rc.lo[0]) + (uint32_t)(rc.hi[1] - rc.lo[1]) + 9 + 3) & ~3u;
    // 32-bit byte offsets when the whole volume spans less than 4 GiB
    const uint64_t esz = vol_desc->dtype == AP_F32 ? 4 : 2;
    const bool bytes32 = !idx64 && (uint64_t)(array3_max_offset(vol_desc) + 1) * esz <= 0xFFFFFFFFull;
    const size_t shm = ((cap * sizeof(float) + 15) & ~size_t(15)) + cap * (bytes32 ? 4 : 8);
    if (shm > 60 * 1024) return AP_ERR_UNSUPPORTED;
    SinoCfgC sc{V, nu, nv, s0, s1, s2, bs[0], bs[1], bs[2], 0u, 0u};
    const int var = ap_tune_variant(AP_TUNE_PROJECT_COLUMN);
    const uint32_t rows = (uint32_t)kColRows[var];
    ColMap cm;
    cm.nkb = (nv + rows - 1) / rows;
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
                    auto launch = [&](auto kern, auto off, uint32_t esz_scale) {
                        using OffT = decltype(off);
                        hipLaunchKernelGGL(kern, grid, dim3(rows), shm, s, rc, sc, cm, mba, ssu, ssv, specs,
                                           (const Tv*)vol, (OffT)(vol_desc->strides[0] * esz_scale),
                                           (OffT)(vol_desc->strides[1] * esz_scale),
                                           (OffT)(vol_desc->strides[2] * esz_scale), (Tp*)sino, (const Tp*)b, cap);
                    };
                    // 64-bit element offsets, or 32-bit byte offsets
                    if (!bytes32) {
                        if (var == 0) launch(k_project_columns<kColRows[0], Tv, Tp, int64_t, false>, int64_t{}, 1);
                        if (var == 1) launch(k_project_columns<kColRows[1], Tv, Tp, int64_t, false>, int64_t{}, 1);
                        if (var == 2) launch(k_project_columns<kColRows[2], Tv, Tp, int64_t, false>, int64_t{}, 1);
                    } else {
                        const uint32_t e = (uint32_t)sizeof(Tv);
                        if (ssu == 1 && ssv == 1) {
                            if (var == 0) launch(k_project_columns<kColRows[0], Tv, Tp, uint32_t, true>, uint32_t{}, e);
                            if (var == 1) launch(k_project_columns<kColRows[1], Tv, Tp, uint32_t, true>, uint32_t{}, e);
                            if (var == 2) launch(k_project_columns<kColRows[2], Tv, Tp, uint32_t, true>, uint32_t{}, e);
                        } else {
                            if (var == 0) launch(k_project_columns<kColRows[0], Tv, Tp, uint32_t, false>, uint32_t{}, e);
                            if (var == 1) launch(k_project_columns<kColRows[1], Tv, Tp, uint32_t, false>, uint32_t{}, e);
                            if (var == 2) launch(k_project_columns<kColRows[2], Tv, Tp, uint32_t, false>, uint32_t{}, e);
                        }
                    }
                    return ap_check_launch("project_columns");
                }
            });
        }
    });
}
