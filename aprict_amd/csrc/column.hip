// column.hip -- column-coherent Siddon forward projector and matched adjoint
// for views whose detector v axis is parallel to z (see column.h).
//
// Forward: a workgroup takes one detector (sub-)column of one view and 128
// consecutive rows (64 or 256 in the other launch-shape variants).  It merges
// the column's x- and y-plane crossings into a segment list in LDS (each
// crossing computes its own rank, so the list is built by all threads at
// once), then every lane walks the list, inserting its own z crossings: per
// voxel a lane does one LDS broadcast read, a few FP32 operations and one
// gather.  Row-adjacent lanes read z-adjacent voxels of the same column, so
// each gather touches one or two cache lines.
//
// Adjoint: a workgroup owns a brick of kTX x kTY x kTZ voxels accumulated in
// LDS over all views as 64-bit fixed point (ds_add_u64).  Per view, each wave
// takes up to kWCols (sub-)columns of the brick's footprint, builds their
// segment lists inside the brick, and every lane scatters weight * y along
// its ray.  Segment boundaries and weights are those of the forward walk.
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
    float unit;        // adjoint fixed-point unit 2^b (adjoint_fixed_point_unit)
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

template <int kRowsF, typename Tv, typename Tp, typename OffT>
__global__ __launch_bounds__(kRowsF) void k_project_columns(
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

// ------------------------------------------------------------------ adjoint
// A workgroup of kAdjT threads owns a brick of kTX x kTY x kTZ voxels
// accumulated in LDS as 64-bit fixed point: float atomic adds in LDS run at a
// small fraction of the rate of integer ones, and integer sums do not depend
// on the order of the adds, so the result is deterministic.  The fixed-point
// scale S is set per workgroup from the largest |y| over the brick's detector
// footprints, so that every term rint(w * y * scale * sec * S) fits in 31 bits
// and the 64-bit sums cannot overflow (adjoint_fixed_point_unit lowers the
// unit below 2^29 for more than 2^33 views x rays per view); the quantisation
// error is about 1e-9 of the largest term.
//
// Each wave processes its own views, one after another, without workgroup
// barriers: it projects the brick's corners to find the detector footprint,
// sets up up to kWCols sub-columns at a time (2-D line, crossing sequences,
// merged segment list inside the brick, all in the wave's own LDS area), then
// every lane takes (sub-column, sub-row) pairs and scatters along its ray.
// Segment boundaries and weights are those of the forward walk.
constexpr int kTX = 16, kTY = 16, kTZ = 32;
constexpr int kAdjT = 1024;
constexpr int kAdjWaves = kAdjT / 64;
constexpr int kWCols = 8;               // sub-columns per wave pass
constexpr int kPass1Loads = 8;          // sinogram loads per lane in flight in the scale pass
// segments of one column inside a brick (<= kTX + kTY - 1) plus 3 padding
// entries, rounded to whole float4 rows
constexpr int kListCap = kTX + kTY + 4;
constexpr int kTXY = kTX > kTY ? kTX : kTY;   // most crossings of one axis in a brick

// What the rows of one sub-column need, and the state that builds its list.
struct ColEntry {
    ColZ z;
    float h0, h1;                  // the column line inside the brick
    int nseg;                      // 0: the column misses the brick
    int j;
    AxisSeq qx, qy;
    int ix0, iy0;
    float aux, auy;
};

struct WaveArea {
    ViewRays vr;                   // the current view (kept here, not in registers)
    ColEntry ce[kWCols];
    float he[kWCols][kListCap];
    int ob[kWCols][kListCap];      // LDS byte offsets of the cells' z lines
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

// Detector footprint [j0, j1] x [k0, k1] of the brick [blo, bhi) in view vr
// (wave-uniform; lanes 0-7 project one corner each).  Returns false when it
// is empty.
// Pixel range of the corners' projections [jmn, jmx] x [kmn, kmx], widened
// by half a pixel and clipped to the detector; the whole detector when a
// corner does not project (bad).
__device__ __forceinline__ bool footprint_range(float jmn, float jmx, float kmn, float kmx, bool bad, uint32_t nu,
                                                uint32_t nv, int& j0, int& j1, int& k0, int& k1) {
    if (bad) {
        j0 = 0; j1 = (int)nu - 1; k0 = 0; k1 = (int)nv - 1;
    } else {
        j0 = max(0, (int)floorf(jmn - 0.501f));
        j1 = min((int)nu - 1, (int)ceilf(jmx + 0.501f));
        k0 = max(0, (int)floorf(kmn - 0.501f));
        k1 = min((int)nv - 1, (int)ceilf(kmx + 0.501f));
    }
    return j0 <= j1 && k0 <= k1;
}

__device__ __forceinline__ void brick_corner(const int blo[3], const int bhi[3], const float cen[3], int c,
                                             float X[3]) {
    X[0] = (float)((c & 4) ? bhi[0] : blo[0]) - cen[0];
    X[1] = (float)((c & 2) ? bhi[1] : blo[1]) - cen[1];
    X[2] = (float)((c & 1) ? bhi[2] : blo[2]) - cen[2];
}

__device__ inline bool brick_footprint(const ViewRays& vr, const int blo[3], const int bhi[3], const float cen[3],
                                       uint32_t nu, uint32_t nv, int lane, int& j0, int& j1, int& k0, int& k1) {
    float jmn = AP_BIG, jmx = -AP_BIG, kmn = AP_BIG, kmx = -AP_BIG, bad = 0.f;
    if (lane < 8) {
        float X[3];
        brick_corner(blo, bhi, cen, lane, X);
        float ju, kv;
        if (proj_pt(vr, X, ju, kv)) {
            jmn = jmx = ju;
            kmn = kmx = kv;
        } else {
            bad = 1.f;
        }
    }
    jmn = lanes8_min(jmn); jmx = lanes8_max(jmx);
    kmn = lanes8_min(kmn); kmx = lanes8_max(kmx);
    return footprint_range(jmn, jmx, kmn, kmx, lanes8_max(bad) > 0.f, nu, nv, j0, j1, k0, k1);
}

// The same footprint found by one lane alone, its 8 corners in turn (for one
// view per lane): the same projections and extremes, hence the same range.
__device__ inline bool lane_footprint(const ViewRays& vr, const int blo[3], const int bhi[3], const float cen[3],
                                      uint32_t nu, uint32_t nv, int& j0, int& j1, int& k0, int& k1) {
    float jmn = AP_BIG, jmx = -AP_BIG, kmn = AP_BIG, kmx = -AP_BIG;
    bool bad = false;
    for (int c = 0; c < 8; c++) {
        float X[3];
        brick_corner(blo, bhi, cen, c, X);
        float ju, kv;
        if (proj_pt(vr, X, ju, kv)) {
            jmn = fminf(jmn, ju); jmx = fmaxf(jmx, ju);
            kmn = fminf(kmn, kv); kmx = fmaxf(kmx, kv);
        } else {
            bad = true;
        }
    }
    return footprint_range(jmn, jmx, kmn, kmx, bad, nu, nv, j0, j1, k0, k1);
}

// floor(x + 0.5) as an integer in one instruction (round half up: the
// quantisation error of a term is uniform in (-1/2, 1/2] of a unit, no bias).
__device__ __forceinline__ int cvt_rpi(float x) {
    int r;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// One fixed-point term: w * wys rounded to the nearest integer (|w * wys| <
// 2^30), sign-extended to 64 bits.
__device__ __forceinline__ unsigned long long fx_term(float w, float wys) {
    return (unsigned long long)(long long)cvt_rpi(w * wys);
}

// A term of the walk.  In a brick where every w * y * scale is >= 0 (the
// footprint's sinogram values and the scale share one sign; kSigned false)
// every rounded term is >= 0, so its 64-bit high word is the constant 0.
// Otherwise (kSigned) the high word is the low word's sign, one more
// instruction per term.
template <bool kSigned>
__device__ __forceinline__ unsigned long long fx_term_walk(float w, float wys) {
    const int lo = cvt_rpi(w * wys);
    const uint32_t hi = kSigned ? (uint32_t)(lo >> 31) : 0u;
    return ((unsigned long long)hi << 32) | (uint32_t)lo;
}

template <typename Tv, typename Tp, typename Tm, typename IdxT>
__global__ __launch_bounds__(kAdjT) void k_backproject_columns(
    RayCfg rc, SinoCfgC sc, float scale, uint32_t ssu, uint32_t ssv,
    const ap_proj_spec* __restrict__ specs, Tv* __restrict__ vol, IdxT vsx, IdxT vsy, IdxT vsz,
    const Tm* __restrict__ mult, IdxT msx, IdxT msy, IdxT msz, const Tp* __restrict__ sino,
    uint32_t nbx, uint32_t nby, uint32_t nbz, uint32_t per_xcd) {
    __shared__ unsigned long long acc[kTX * kTY * kTZ];
    __shared__ __align__(16) WaveArea wa_all[kAdjWaves];
    __shared__ float wmin[kAdjWaves], wmax[kAdjWaves];
    const uint32_t L = blockIdx.x;
    const uint32_t logical = (L % kXcdC) * per_xcd + L / kXcdC;
    if (logical >= nbx * nby * nbz) return;
    const uint32_t bk = logical % nbz, bj = (logical / nbz) % nby, bi = logical / (nbz * nby);
    const int blo[3] = {rc.lo[0] + (int)bi * kTX, rc.lo[1] + (int)bj * kTY, rc.lo[2] + (int)bk * kTZ};
    const int bhi[3] = {min(blo[0] + kTX, rc.hi[0]), min(blo[1] + kTY, rc.hi[1]), min(blo[2] + kTZ, rc.hi[2])};
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    WaveArea& W = wa_all[wave];
    const float cen[3] = {0.5f * (float)(rc.lo[0] + rc.hi[0]), 0.5f * (float)(rc.lo[1] + rc.hi[1]),
                          0.5f * (float)(rc.lo[2] + rc.hi[2])};
    // pass 1: the range [ylo, yhi] of the sinogram values over the brick's
    // footprints sets the fixed-point scale (largest |y|) and tells whether
    // every term has one sign.  Wave w takes views w, w + 16, ...: its lanes
    // first find the footprints of 64 of them at once (lane l: view
    // pb + 16 l), then the wave reads each footprint's values together
    float ylo = 0.f, yhi = 0.f;
    for (uint32_t pb = wave; pb < sc.V; pb += 64u * kAdjWaves) {
        const uint32_t pl = pb + (uint32_t)lane * kAdjWaves;
        int fj0 = 0, fj1 = -1, fk0 = 0, fk1 = -1;   // empty: not upright or beyond the last view
        if (pl < sc.V && view_is_upright(specs, pl))
            lane_footprint(make_view_rays(rc, specs, pl), blo, bhi, cen, sc.nu, sc.nv, fj0, fj1, fk0, fk1);
        const int nq = (int)min(64u, (sc.V - pb + kAdjWaves - 1) / kAdjWaves);
        for (int iq = 0; iq < nq; iq++) {
            const uint32_t p = pb + (uint32_t)iq * kAdjWaves;
            const int j0 = __builtin_amdgcn_readlane(fj0, iq), j1 = __builtin_amdgcn_readlane(fj1, iq);
            const int k0 = __builtin_amdgcn_readlane(fk0, iq), k1 = __builtin_amdgcn_readlane(fk1, iq);
            if (j0 > j1 || k0 > k1) continue;   // wave-uniform
            if (sc.vec16) {
                // rows [k0, k1] widened to aligned 16-byte groups of kVec values
                // (nv % kVec == 0 keeps them inside the detector column); the
                // extra rows only loosen the bound
                constexpr int kVec = 16 / (int)sizeof(Tp);
                const int ka = k0 & ~(kVec - 1), ng = ((k1 | (kVec - 1)) - ka + 1) / kVec;
                const int ngrp = (j1 - j0 + 1) * ng;
                const UDiv dg = make_udiv(ng);
                // kPass1Loads loads per lane in flight before any is reduced;
                // indices past the footprint are clamped to its last group (a
                // repeated group leaves the maximum unchanged)
                for (int e0 = lane; e0 < ngrp; e0 += 64 * kPass1Loads) {
                    uint4 raw[kPass1Loads];
    #pragma unroll
                    for (int u = 0; u < kPass1Loads; u++) {
                        const int e = min(e0 + 64 * u, ngrp - 1);
                        const int jj = udiv(e, dg);
                        raw[u] = *reinterpret_cast<const uint4*>(
                            sino + sino_off(sc, p, j0 + jj, ka + kVec * (e - jj * ng)));
                    }
    #pragma unroll
                    for (int u = 0; u < kPass1Loads; u++) {
                        const Tp* v = reinterpret_cast<const Tp*>(&raw[u]);
                        float mn = 0.f, mx = 0.f, sum = 0.f;
    #pragma unroll
                        for (int q = 0; q < kVec; q++) {
                            const float y = to_f32(v[q]);
                            mn = fminf(mn, y);
                            mx = fmaxf(mx, y);
                            sum += y;
                        }
                        ylo = fminf(ylo, mn);
                        yhi = fmaxf(yhi, mx);
                        if (sum != sum) yhi = __builtin_inff();   // NaN (or opposite infinities)
                    }
                }
                continue;
            }
            const int nk = k1 - k0 + 1, npix = (j1 - j0 + 1) * nk;
            const UDiv dk = make_udiv(nk);
            for (int e0 = lane; e0 < npix; e0 += 64 * kPass1Loads) {
                float yv[kPass1Loads];
    #pragma unroll
                for (int u = 0; u < kPass1Loads; u++) {
                    const int e = min(e0 + 64 * u, npix - 1);
                    const int jj = udiv(e, dk);
                    yv[u] = to_f32(sino[sino_off(sc, p, j0 + jj, k0 + e - jj * nk)]);
                }
    #pragma unroll
                for (int u = 0; u < kPass1Loads; u++) {
                    ylo = fminf(ylo, yv[u]);   // a NaN sample is caught below
                    yhi = fmaxf(yhi, yv[u]);
                    if (yv[u] != yv[u]) yhi = __builtin_inff();
                }
            }
        }
    }
    ylo = wave_min(ylo);
    yhi = wave_max(yhi);
    if (lane == 0) {
        wmin[wave] = ylo;
        wmax[wave] = yhi;
    }
    for (int e = tid; e < kTX * kTY * kTZ; e += kAdjT) acc[e] = 0ull;
    __syncthreads();
    ylo = 0.f;
    yhi = 0.f;
    for (int w = 0; w < kAdjWaves; w++) {
        ylo = fminf(ylo, wmin[w]);
        yhi = fmaxf(yhi, wmax[w]);
    }
    const float amax = fmaxf(yhi, -ylo) * fabsf(scale);
    // every w * y * scale >= 0 (or <= 0 with the sign folded into the scale):
    // the walk's terms need no sign extension
    const bool one_sign = scale >= 0.f ? ylo >= 0.f : yhi <= 0.f;
    if (amax == 0.f) return;   // block-uniform: nothing to add
    if (!(amax < __builtin_inff())) {
        // non-finite data: the sum is not representable; propagate NaN
        for (int e = tid; e < kTX * kTY * kTZ; e += kAdjT) {
            const int li = e / (kTY * kTZ), lj = (e / kTZ) % kTY, lk = e % kTZ;
            const int gi = blo[0] + li, gj = blo[1] + lj, gk = blo[2] + lk;
            if (gi >= bhi[0] || gj >= bhi[1] || gk >= bhi[2]) continue;
            vol[(IdxT)gi * vsx + (IdxT)gj * vsy + (IdxT)gk * vsz] = from_f32<Tv>(__builtin_nanf(""));
        }
        return;
    }
    // |w * sec| <= sqrt(3) < 2, so every term is below 2 * unit in magnitude
    const float S = sc.unit / amax;
    const float scaleS = scale * S;
    // the views, with or without sign extension of the terms (block-uniform)
    auto all_views = [&](auto sign_kind) {
    constexpr bool kSigned = decltype(sign_kind)::value;
    for (uint32_t p = wave; p < sc.V; p += kAdjWaves) {
        if (!view_is_upright(specs, p)) continue;   // wave-uniform
        wave_sync_lds();   // the previous view is done with the area
        if (lane == 0) W.vr = make_view_rays(rc, specs, p);
        wave_sync_lds();
        int j0, j1, k0, k1;
        if (!brick_footprint(W.vr, blo, bhi, cen, sc.nu, sc.nv, lane, j0, j1, k0, k1)) continue;
        const int ncol = (j1 - j0 + 1) * (int)ssu;
        const int nrow = (k1 - k0 + 1) * (int)ssv;
        for (int c0 = 0; c0 < ncol; c0 += kWCols) {
            const int nc = min(kWCols, ncol - c0);
            wave_sync_lds();   // the previous pass is done with the tables
            if (lane < nc) {
                ColEntry& E = W.ce[lane];
                const int g = c0 + lane;
                E.j = j0 + g / (int)ssu;
                E.nseg = 0;
                E.qx.n = 0;   // the list builder reads the counts of every column
                E.qy.n = 0;
                Col2D C;
                if (col_setup(W.vr, rc.lo, rc.hi, (float)E.j + sub_off((uint32_t)(g % (int)ssu), ssu), C)) {
                    float h0, h1;
                    col_box(C, blo, bhi, h0, h1);
                    if (h0 < h1) {
                        E.z = C.z;
                        E.h0 = h0;
                        E.h1 = h1;
                        E.ix0 = axis_index_at(C.r[0], C.c[0], C.P0[0], C.u[0], C.sg[0], C.ifix[0], h0, blo[0],
                                              bhi[0]);
                        E.iy0 = axis_index_at(C.r[1], C.c[1], C.P0[1], C.u[1], C.sg[1], C.ifix[1], h0, blo[1],
                                              bhi[1]);
                        E.qx = axis_seq(C, 0, E.ix0, h1, blo[0], bhi[0]);
                        E.qy = axis_seq(C, 1, E.iy0, h1, blo[1], bhi[1]);
                        E.aux = fabsf(C.u[0]);
                        E.auy = fabsf(C.u[1]);
                        E.nseg = E.qx.n + E.qy.n + 1;
                        W.ob[lane][0] = ((E.ix0 - blo[0]) * kTY + (E.iy0 - blo[1])) * kTZ * 8;
                        // the last segment ends at h1; 3 zero-length padding
                        // segments repeat the final cell for the 4-wide walk
                        const int ixe = min(max(E.ix0 + E.qx.sg * E.qx.n, blo[0]), bhi[0] - 1);
                        const int iye = min(max(E.iy0 + E.qy.sg * E.qy.n, blo[1]), bhi[1] - 1);
                        const int oe = ((ixe - blo[0]) * kTY + (iye - blo[1])) * kTZ * 8;
                        for (int q = 0; q < 4; q++) W.he[lane][E.nseg - 1 + q] = h1;
                        for (int q = 1; q < 4; q++) W.ob[lane][E.nseg - 1 + q] = oe;
                    }
                }
            }
            wave_sync_lds();
            // merged x-y crossing lists: (column, crossing) pairs across the wave
            for (int e = lane; e < nc * kTXY; e += 64) {
                const int g = e / kTXY, m = e % kTXY;
                const ColEntry& E = W.ce[g];
                if (m < E.qx.n) {
                    const float h = seq_at(E.qx, m);
                    const int cy = seq_count(E.qy, h, true, E.auy);
                    const int ix = min(max(E.ix0 + E.qx.sg * (m + 1), blo[0]), bhi[0] - 1);
                    const int iy = min(max(E.iy0 + E.qy.sg * cy, blo[1]), bhi[1] - 1);
                    W.he[g][m + cy] = h;
                    W.ob[g][m + cy + 1] = ((ix - blo[0]) * kTY + (iy - blo[1])) * kTZ * 8;
                }
                if (m < E.qy.n) {
                    const float h = seq_at(E.qy, m);
                    const int cx = seq_count(E.qx, h, false, E.aux);
                    const int ix = min(max(E.ix0 + E.qx.sg * cx, blo[0]), bhi[0] - 1);
                    const int iy = min(max(E.iy0 + E.qy.sg * (m + 1), blo[1]), bhi[1] - 1);
                    W.he[g][m + cx] = h;
                    W.ob[g][m + cx + 1] = ((ix - blo[0]) * kTY + (iy - blo[1])) * kTZ * 8;
                }
            }
            wave_sync_lds();
            // (sub-column, sub-row) pairs: every lane scatters one ray at a time;
            // the sinogram value of the lane's next pair is fetched one pair ahead
            const int npair = nc * nrow;
            const UDiv drow = make_udiv(nrow), dsv = make_udiv((int)ssv);
            // one sub-row per pixel (ssv == 1, wave-uniform): the sub-row is the
            // row and its offset is 0 (what sub_off(0, 1) gives), no divisions
            const bool ss1 = ssv == 1;
            auto yval = [&](int e0) -> float {
                const int e = min(e0, npair - 1);
                const int g = udiv(e, drow);
                const int k = k0 + (ss1 ? e - g * nrow : udiv(e - g * nrow, dsv));
                return to_f32(sino[sino_off(sc, p, W.ce[g].j, k)]);
            };
            auto pair = [&](int e, float y) {
                const int g = udiv(e, drow), kk = e - g * nrow;
                const ColEntry& E = W.ce[g];
                if (E.nseg == 0 || y == 0.f) return;
                const int kq = ss1 ? kk : udiv(kk, dsv);
                const int k = k0 + kq;
                const float so = ss1 ? 0.f : sub_off((uint32_t)(kk - kq * (int)ssv), ssv);
                LaneZ Z;
                if (!lane_setup(E.z, (float)k + so, rc.lo[2], rc.hi[2], Z)) return;
                float zin, zout;
                lane_box(Z, blo[2], bhi[2], zin, zout);
                const float t0 = fmaxf(E.h0, zin), t1 = fminf(E.h1, zout);
                if (!(t0 < t1)) return;
                const int iz = axis_index_at(Z.r, Z.c, Z.P0, Z.u, Z.sg, Z.ifix, t0, blo[2], bhi[2]);
                float pz = Z.sg > 0 ? (float)(iz + 1) - Z.P0 : (float)iz - Z.P0;
                float hz = Z.sg == 0 ? AP_BIG : tcross(pz, Z.r, Z.c);
                const float gz = (float)Z.sg;
                int zl = iz - blo[2];
                const float wys = y * scaleS * Z.sec;
                const float* he = W.he[g];
                const int* ob = W.ob[g];
                float hp = t0;
                if (Z.u > -kSingleSlope && Z.u < kSingleSlope) {
                    // at most one z crossing per segment: 4 segments per step
                    // (LDS byte offsets), ending with the chunk that reaches t1
                    char* accb = reinterpret_cast<char*>(acc);
                    int zb = zl * 8;
                    const int dzb = Z.sg * 8;
                    // the list of chunk i+1 is read before the adds of chunk i
                    // issue, so its wait does not include them (LDS returns in
                    // order); chunks start at most at kListCap - 8, so the read
                    // ahead stays inside the list; the padding ends the walk
                    // with the chunk whose last entry reaches t1.  The walk
                    // starts at segment 0: segments that end before t0 clamp
                    // to [t0, t0], weight 0 and hold no z crossing (the next
                    // one, hz, lies beyond t0), so they add zero terms to
                    // cells of the brick;
                    // most rows enter the brick at h0, where the walk starts
                    // anyway, and a search for the first segment would cost
                    // more than the few extra chunks of the others
                    int s = 0;
                    float4 h4 = *reinterpret_cast<const float4*>(he + s);
                    int4 o4 = *reinterpret_cast<const int4*>(ob + s);
                    // complete the first read here, so that the loop head
                    // never waits for outstanding adds
                    asm volatile("" : "+v"(h4.x), "+v"(h4.y), "+v"(h4.z), "+v"(h4.w), "+v"(o4.x), "+v"(o4.y),
                                 "+v"(o4.z), "+v"(o4.w));
                    // one chunk of 4 segments: weights and addresses first, then
                    // the adds, so no LDS read is outstanding while they issue.
                    // A lane adds the part after a z crossing only where it
                    // crossed (about one segment in 1/|slope|): the LDS
                    // atomics, not the arithmetic, bound the walk, and the
                    // masked add costs less.  The part before is added
                    // unmasked; it is rarely empty, and a mask there costs
                    // more in branches and waits than it saves
                    auto chunk = [&](const float4& hc4, const int4& oc4) {
                        const float hv[4] = {hc4.x, hc4.y, hc4.z, hc4.w};
                        const int ov[4] = {oc4.x, oc4.y, oc4.z, oc4.w};
                        float wb[4], wa[4];
                        int a0[4], a1[4];
                        bool crs[4];
#pragma unroll
                        for (int q = 0; q < 4; q++) {
                            const float ce = __builtin_amdgcn_fmed3f(hv[q], t0, t1);
                            const bool cr = hz < ce;
                            crs[q] = cr;
                            const float c = cr ? hz : ce;
                            wb[q] = c - hp;
                            wa[q] = ce - c;
                            hp = ce;
                            // z step of this segment; kept opaque so that the
                            // two addresses and the next offset take one add
                            // each (not an add and a select per address)
                            int d = cr ? dzb : 0;
                            asm volatile("" : "+v"(d));
                            a0[q] = ov[q] + zb;
                            a1[q] = a0[q] + d;
                            zb += d;
                            pz += cr ? gz : 0.f;
                            hz = tcross(pz, Z.r, Z.c);
                        }
                        // keep the compiler from sinking the address arithmetic
                        // (which waits on the list read) between the adds
                        asm volatile("" : "+v"(a0[0]), "+v"(a0[1]), "+v"(a0[2]), "+v"(a0[3]), "+v"(a1[0]),
                                     "+v"(a1[1]), "+v"(a1[2]), "+v"(a1[3]));
#pragma unroll
                        for (int q = 0; q < 4; q++) {
                            atomicAdd(reinterpret_cast<unsigned long long*>(accb + a0[q]),
                                      fx_term_walk<kSigned>(wb[q], wys));
                            if (crs[q])
                                atomicAdd(reinterpret_cast<unsigned long long*>(accb + a1[q]),
                                          fx_term_walk<kSigned>(wa[q], wys));
                        }
                    };
                    // two chunks per turn with alternating list registers
                    for (;;) {
                        const float4 h4b = *reinterpret_cast<const float4*>(he + s + 4);
                        const int4 o4b = *reinterpret_cast<const int4*>(ob + s + 4);
                        chunk(h4, o4);
                        if (h4.w >= t1) break;
                        s += 4;
                        h4 = *reinterpret_cast<const float4*>(he + s + 4);
                        o4 = *reinterpret_cast<const int4*>(ob + s + 4);
                        chunk(h4b, o4b);
                        if (h4b.w >= t1) break;
                        s += 4;
                    }
                } else {
                    // first segment ending after t0 (he[nseg-1] = h1 > t0)
                    const int s0 = first_above(he, E.nseg, t0);
                    for (int s = s0; s < E.nseg; s++) {
                        const float cend = __builtin_amdgcn_fmed3f(he[s], t0, t1);
                        const int o = ob[s] / 8;
                        while (hz < cend) {
                            atomicAdd(&acc[o + zl], fx_term(hz - hp, wys));
                            hp = hz;
                            pz += gz;
                            hz = tcross(pz, Z.r, Z.c);
                            zl += Z.sg;
                        }
                        atomicAdd(&acc[o + zl], fx_term(cend - hp, wys));
                        hp = cend;
                        if (cend >= t1) break;
                    }
                }
            };
            // two pairs per turn with alternating registers for the value
            // fetched ahead (the load is unconditional, from a clamped index),
            // so no register copy or branch makes a use wait for the newer load
            float ya = yval(lane);
            for (int e = lane; e < npair; e += 128) {
                const float yb = yval(e + 64);
                pair(e, ya);
                if (e + 64 >= npair) break;
                ya = yval(e + 128);
                pair(e + 64, yb);
            }
        }
    }
    };
    if (one_sign)
        all_views(std::false_type{});
    else
        all_views(std::true_type{});
    __syncthreads();
    const double inv_s = 1.0 / (double)S;
    for (int e = tid; e < kTX * kTY * kTZ; e += kAdjT) {
        const int li = e / (kTY * kTZ), lj = (e / kTZ) % kTY, lk = e % kTZ;
        const int gi = blo[0] + li, gj = blo[1] + lj, gk = blo[2] + lk;
        if (gi >= bhi[0] || gj >= bhi[1] || gk >= bhi[2]) continue;
        const long long a = (long long)acc[e];
        if (a == 0) continue;
        const IdxT ov = (IdxT)gi * vsx + (IdxT)gj * vsy + (IdxT)gk * vsz;
        const float m = mult ? to_f32(mult[(IdxT)gi * msx + (IdxT)gj * msy + (IdxT)gk * msz]) : 1.f;
        vol[ov] = from_f32<Tv>(__builtin_fmaf(m, (float)((double)a * inv_s), to_f32(vol[ov])));
    }
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
    SinoCfgC sc{V, nu, nv, s0, s1, s2, bs[0], bs[1], bs[2], 0u, 0.f};
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
                        if (var == 0) launch(k_project_columns<kColRows[0], Tv, Tp, int64_t>, int64_t{}, 1);
                        if (var == 1) launch(k_project_columns<kColRows[1], Tv, Tp, int64_t>, int64_t{}, 1);
                        if (var == 2) launch(k_project_columns<kColRows[2], Tv, Tp, int64_t>, int64_t{}, 1);
                    } else {
                        const uint32_t e = (uint32_t)sizeof(Tv);
                        if (var == 0) launch(k_project_columns<kColRows[0], Tv, Tp, uint32_t>, uint32_t{}, e);
                        if (var == 1) launch(k_project_columns<kColRows[1], Tv, Tp, uint32_t>, uint32_t{}, e);
                        if (var == 2) launch(k_project_columns<kColRows[2], Tv, Tp, uint32_t>, uint32_t{}, e);
                    }
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
    SinoCfgC sc{V, nu, nv, s0, s1, s2, s0, s1, s2, 0u};
    sc.o32 = (double)(V ? V - 1 : 0) * (double)s0 + (double)(nu ? nu - 1 : 0) * (double)s1 +
                     (double)(nv ? nv - 1 : 0) * (double)s2 < 2147483647.0;
    {
        const int64_t kv = sino_dtype == AP_F32 ? 4 : sino_dtype == AP_F16 ? 8 : 0;
        sc.vec16 = kv > 0 && s2 == 1 && nv % kv == 0 && s1 % kv == 0 && s0 % kv == 0 &&
                   ((uintptr_t)sino & 15) == 0;
    }
    sc.unit = adjoint_fixed_point_unit((double)V * nu * nv * ssu * ssv);
    if (sc.unit == 0.f) return ap_fail(AP_ERR_INVALID, "too many views x rays per view for the adjoint");
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
