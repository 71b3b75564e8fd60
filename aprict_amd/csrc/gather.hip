// gather.hip -- matched adjoint of the Siddon forward projector for views
// whose detector v axis is parallel to z and whose u axis has no z component
// (circular and helical scans with upright panels), computed as a gather over
// voxels: no atomics, every voxel sum accumulated in FP64 registers.
//
// Geometry.  For such a view all rays through detector (sub-)column q lie in
// one vertical plane through the source: they share their x-y line, and
// along that line, parametrised by the horizontal distance h from the source,
// ray t has height z_t(h) = sz + h * dz_t / l, where dz_t is the height of
// its detector point above the source and l the horizontal source-pixel
// distance.  Order the rays by height (dz_t = dz0 + t dzs, dzs > 0).  The
// column's line crosses x-y cell (i, j) over [h_a, h_b]; the Siddon length of
// ray t in voxel (i, j, k) is sec_t times the measure of the h in [h_a, h_b]
// where k <= z_t(h) < k + 1.  With
//     F_t(k) = |{h in [h_a, h_b] : z_t(h) < k}|   and
//     G(k)   = sum_t yhat_t F_t(k),   yhat_t = y_t sec_t,
// the adjoint's contribution of the column to voxel k is G(k+1) - G(k).
// Rays entirely below plane k contribute h_b - h_a each, so
//     G(k) = (h_b - h_a) P[tf] + F_tf(k) yhat_tf,
// where P is the exclusive prefix sum of yhat over t and tf the first ray
// that is not entirely below k.  The next rays are entirely above k when the
// rays of a column are further apart in z than one of them climbs within a
// cell (kappa below); (cell, column) pairs where that is not certain add
// F_t(k) yhat_t of every ray that may straddle the plane.  tf is the integer
// part of a linear function kappa of k, P is read from LDS, and F_tf(k)
// follows from the crossing of plane k by ray tf.
//
// A pre-pass (k_gather_prep) writes, per view and sub-column, the FP64
// prefix sums P of the sec-weighted sinogram column (in z order), the
// column's x-y crossing constants (the forward's col_setup, so the cell
// boundaries along h are the forward's own FP32 values) and, per view, the
// reciprocal heights 1/dz_t and the number t0 of rows whose ray does not
// rise.  The main kernel (k_gather_adjoint) gives each workgroup (16 waves)
// a tile of 8x8 x-y cells and a z range of 64*kJ boundaries; each wave owns
// kC cells, each lane one boundary per 64 in z.  The work comes in items, a
// view and a group of the columns crossing the tile: while the workgroup
// works on one item, the next one's prefix window is loaded straight into
// the other of two LDS buffers (16-byte loads).  Per item each wave derives
// its (cell, column) pairs one per lane, then for every pair every lane
// evaluates G at its boundaries (10 VALU operations and two LDS reads on
// tiles that lie on one side of the source plane, 16 otherwise) and adds it
// to its FP64 accumulators.  The part of ray tf below plane k is one FP64 FMA
// in units of the cell's chord, clamped to [0, 1] by the FMA's clamp modifier.  The voxel values are differences of
// neighbouring boundaries, scaled and added to the volume once at the end.
// Tiles run in blocks of 4x8, so that the workgroups on one XCD at one time
// share their views' prefix windows through its L2.
//
// Precision: the sums P and the accumulators are FP64, so a voxel's value is
// exact to the FP32 rounding of its own terms whatever the dynamic range of
// the sinogram around it; the crossing of plane k by ray tf is formed in FP64
// from the column's closest approach, with l / dh rounded to FP32 (the FP32
// error of the forward's own crossing parameters).  A ray lying in a z plane (the forward's horizontal
// rows, |slope| < 1e-20) belongs to the voxel below the plane, as in the
// forward; such rays are kept out of P and added separately.
#include <mutex>

#include "column.h"
#include "gather.h"

namespace {

constexpr int kXcdG = 8;
constexpr int kGWin = 57344;       // bytes of an LDS window buffer (56 KB; two per workgroup)
constexpr int kGMaxCols = 16;      // columns staged at once
constexpr int kGRowMax = 2048;     // rows of a window (at most)

// per-(view, sub-column) record
struct GCol {
    float r[2], c[2], P0[2];   // x-y crossing parameters along h - hc (col_setup)
    int32_t sg[2], ifix[2];
    float hs, hd;              // source and pixel along h - hc (hd <= hs: no ray meets the chunk)
    float l, hc;               // horizontal source-pixel distance, closest approach to the chunk centre
    double gam;                // hc / l
    float yh;                  // value of the column's rays lying in a z plane (sec = 1), 0 if none
    int32_t ifixh;             // the voxel row they belong to
    float lods;                // l / dzs
    float pad;
};
static_assert(sizeof(GCol) == 80, "GCol layout");

__host__ __device__ inline size_t g_align(size_t x) { return (x + 255) & ~size_t(255); }

// Rows past the last ray are padded (P: the total, 1/dz: 0) far enough that
// the window loads (whole 256-byte chunks from a row <= NT + 1) stay inside
// the row: kGPad >= 32 + 2.
constexpr uint32_t kGPad = 40;

struct GLayout {
    size_t views, count, cols, pref, iz, total;
    uint32_t cs;   // entries per column of P and per view of 1/dz: NT + kGPad
};

__host__ __device__ inline GLayout g_layout(uint32_t V, uint32_t NQ, uint32_t NT) {
    GLayout L;
    L.cs = NT + kGPad;
    L.views = 0;
    L.count = g_align((size_t)V * sizeof(GView));
    L.cols = L.count + 256;
    // one spare record: the record loads copy whole 256-byte chunks
    L.pref = L.cols + g_align(((size_t)V * NQ + 4) * sizeof(GCol));
    L.iz = L.pref + g_align((size_t)V * NQ * L.cs * sizeof(double));
    L.total = L.iz + g_align((size_t)V * L.cs * sizeof(double));
    return L;
}

struct GCfg {
    uint32_t V, nu, nv, ssu, ssv, NQ, NT, cs;
    int64_t s0, s1, s2;   // sinogram element strides
};

// ---------------------------------------------------------------------------
// pre-pass: one wave per (view, sub-column)
// ---------------------------------------------------------------------------
template <typename Tp>
__global__ __launch_bounds__(256) void k_gather_prep(RayCfg rc, GCfg gc, const ap_proj_spec* __restrict__ specs,
                                                     const Tp* __restrict__ sino, GView* __restrict__ gviews,
                                                     GCol* __restrict__ gcols, double* __restrict__ pref,
                                                     double* __restrict__ izt) {
    const int lane = threadIdx.x & 63;
    const uint64_t w = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= (uint64_t)gc.V * gc.NQ) return;
    const uint32_t p = (uint32_t)(w / gc.NQ), q = (uint32_t)(w % gc.NQ);
    GView gv;
    const bool ok = gather_view(rc, specs, p, gc.ssu, gc.ssv, gc.NT, gv);
    if (q == 0 && lane == 0) gviews[p] = gv;
    if (!ok) return;
    const ViewRays vr = make_view_rays(rc, specs, p);
    const uint32_t j = q / gc.ssu;
    const float ju = (float)j + sub_offset_g(q % gc.ssu, gc.ssu);
    Col2D C;
    const bool hit = col_setup(vr, rc.lo, rc.hi, ju, C);
    double* P = pref + (size_t)w * gc.cs;
    const uint32_t NT = gc.NT;
    double carry = 0.0, yh = 0.0;
    int ifixh = INT32_MIN;
    // rows in z order, 64 at a time (coalesced): inclusive wave scan plus the
    // running total of the previous chunks; P[t] is the exclusive prefix
    for (uint32_t t0 = 0; t0 < NT + kGPad; t0 += 64) {
        const uint32_t t = t0 + (uint32_t)lane;
        double yv = 0.0;
        if (hit && t < NT) {
            const uint32_t sr = gv.flip ? NT - 1 - t : t;
            const uint32_t k = sr / gc.ssv;
            const float kv = (float)k + sub_offset_g(sr % gc.ssv, gc.ssv);
            const double m = fma((double)kv, C.z.dm, C.z.m0);
            const float mf = (float)m;
            const float y = to_f32(sino[(int64_t)p * gc.s0 + (int64_t)j * gc.s1 + (int64_t)k * gc.s2]);
            if (fabsf(mf) < 1e-20f) {
                // lies in its z plane: the forward's horizontal row (lane_setup)
                yh += (double)y;
                ifixh = (int)ceil(fma(C.z.hc, m, C.z.zc)) - 1;
            } else {
                yv = (double)(y * __builtin_amdgcn_sqrtf(__builtin_fmaf(mf, mf, 1.f)));
            }
        }
        double incl = yv;
        for (int o = 1; o < 64; o <<= 1) {
            const double v = __shfl_up(incl, o, 64);
            if (lane >= o) incl += v;
        }
        if (t < gc.cs) P[t] = carry + incl - yv;
        carry += __shfl(incl, 63, 64);
    }
    // the horizontal rows of the column (at most one in practice; summed)
    for (int o = 32; o > 0; o >>= 1) {
        yh += __shfl_xor(yh, o, 64);
        ifixh = max(ifixh, __shfl_xor(ifixh, o, 64));
    }
    if (lane == 0) {
        GCol g;
        for (int a = 0; a < 2; a++) {
            g.r[a] = hit ? C.r[a] : 0.f;
            g.c[a] = hit ? C.c[a] : 0.f;
            g.P0[a] = hit ? C.P0[a] : 0.f;
            g.sg[a] = hit ? C.sg[a] : 0;
            g.ifix[a] = hit ? C.ifix[a] : 0;
        }
        g.hs = hit ? C.hs : 0.f;
        g.hd = hit ? C.hd : 0.f;   // hd <= hs: empty
        const double qx = vr.Q00[0] + (double)ju * vr.U[0] - vr.S[0];
        const double qy = vr.Q00[1] + (double)ju * vr.U[1] - vr.S[1];
        const double l = hit ? sqrt(qx * qx + qy * qy) : 1.0;
        g.l = (float)l;
        g.hc = hit ? (float)C.z.hc : 0.f;
        g.gam = hit ? C.z.hc / l : 0.0;
        g.yh = (float)yh;
        g.ifixh = ifixh == INT32_MIN ? 0 : ifixh;
        g.lods = (float)(l / (double)gv.dzs);
        g.pad = 0.f;
        gcols[w] = g;
    }
    // 1/dz per row (0 past the last ray), once per view; and the number of
    // rows whose ray does not rise (dz <= 0: the rows before the first rising
    // one, since dz grows along the z order)
    if (q == 0) {
        uint32_t nfall = 0;
        for (uint32_t t = (uint32_t)lane; t < gc.cs; t += 64) {
            double v = 0.0;
            bool fall = false;
            if (t < NT) {
                const uint32_t sr = gv.flip ? NT - 1 - t : t;
                const float kv = (float)(sr / gc.ssv) + sub_offset_g(sr % gc.ssv, gc.ssv);
                const double dz = vr.Q00[2] + (double)kv * vr.V[2] - vr.S[2];
                v = dz != 0.0 ? 1.0 / dz : 0.0;
                fall = !(dz > 0.0);
            }
            izt[(size_t)p * gc.cs + t] = v;
            nfall += (uint32_t)__popcll(__ballot(fall));
        }
        if (lane == 0) gviews[p].t0 = (int32_t)nfall;
    }
}

// number of views the gather kernel does not take (for the general adjoint's early exit)
__global__ __launch_bounds__(256) void k_gather_count(const GView* __restrict__ gviews, uint32_t V,
                                                      uint32_t* __restrict__ left) {
    __shared__ uint32_t part[4];
    uint32_t n = 0;
    for (uint32_t p = threadIdx.x; p < V; p += 256) n += gviews[p].ok ? 0u : 1u;
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) *left = part[0] + part[1] + part[2] + part[3];
}

// ---------------------------------------------------------------------------
// main kernel
// ---------------------------------------------------------------------------
// Tiles are taken in blocks of kSTX x kSTY (x, y) tiles of one z layer: the
// workgroups that run at the same time on one XCD cover a compact block of
// cells, whose views' prefix windows they share through that XCD's L2.
constexpr uint32_t kSTX = 4, kSTY = 8, kST = kSTX * kSTY;

struct GTile {
    uint32_t ntx, nty, ntz;   // tiles along x, y, z
    uint32_t nsx, nsy;        // tile blocks along x, y
    uint32_t total, per_xcd;  // block slots (ntz nsx nsy kST), slots per XCD
};

// integer part of x, one instruction (x within int range)
__device__ __forceinline__ int cvt_flr(float x) {
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// prefix window of one view for the tile (views taken 64 at a time, one per lane)
struct GWin {
    int qlo, qhi, tlo, W;   // columns [qlo, qhi], rows tlo .. tlo + W - 1 of P (qlo > qhi: nothing)
    float kc;                // kappa offset: 1 - tlo - dz0 / dzs
    int ncg;                 // columns per staging group
    int pad[2];
};

// a view's parameters, as wave-uniform values
struct GVu {
    float sz, sx, sy;
    int t0;
    float ax, ay, bx, by;
    float szlo;   // the source height is sz + szlo (FP64 crossings use both)
};

template <typename Tv, typename Tm, int kW, int kC, int kJ, int kTX>
__global__ __launch_bounds__(kW * 64) void k_gather_adjoint(
    RayCfg rc, GCfg gc, GTile tl, const GView* __restrict__ gviews, const GCol* __restrict__ gcols,
    const double* __restrict__ pref, const double* __restrict__ izt, double scale, Tv* __restrict__ vol,
    int64_t vsx, int64_t vsy, int64_t vsz, const Tm* __restrict__ mult, int64_t msx, int64_t msy, int64_t msz) {
    constexpr int kCells = kW * kC;
    constexpr int kTY = kCells / kTX;
    constexpr int kNQc = (kTX + 1) * (kTY + 1);
    static_assert(kTX * kTY == kCells, "tile shape");
    static_assert(64 / kC >= kGMaxCols, "one setup pass per group");
    // Two staging buffers: while the workgroup works on one (view, column group)
    // item from one buffer, the next item's window is loaded into the other.
    // A buffer holds the item's window -- rows tlo .. tlo + Wp - 1 of 1/dz, then
    // the same rows of P for each column of the group --, the group's column
    // records and the projections of the tile's cell corners.
    struct GBuf {
        double win[kGWin / 8];
        GCol cols[kGMaxCols];
        float qc[(kNQc + 3) & ~3];
    };
    __shared__ __align__(16) GBuf Buf0;
    __shared__ __align__(16) GBuf Buf1;
    __shared__ GWin Win[64];
    // the batch's view parameters, one 32-byte record per view: sz, sx, sy, t0 (bits),
    // ax, ay, bx, by -- read from LDS at the start of a step, so that no step waits
    // on a global load of its view record
    __shared__ __align__(16) float Vw[64 * 8];
    __shared__ float VwLo[64];   // the views' szlo
    // per-wave (cell, column) pair records of the setup pass, one per lane:
    // l / dh, dh = h_b - h_a, alpha_a, alpha_b, 0, window offset | flags, H (FP64).
    // The records start 8 bytes past a 16-byte boundary: the compiler then reads
    // them as 8-byte pairs, which keep its knowledge that they do not touch the
    // window loads in flight (its 16-byte reads lose it and wait for them)
    __shared__ __align__(16) float PrecA[kW * 64 * 8 + 2];
    auto prec = [&](int w, int l) -> float* { return PrecA + 2 + (w * 64 + l) * 8; };
    const uint32_t L = blockIdx.x;
    const uint32_t logical = (L % kXcdG) * tl.per_xcd + L / kXcdG;
    if (logical >= tl.total) return;
    const uint32_t per_layer = tl.nsx * tl.nsy * kST;
    const uint32_t tz = logical / per_layer, r = logical % per_layer;
    const uint32_t st = r / kST, w = r % kST;
    const uint32_t tx = (st % tl.nsx) * kSTX + w % kSTX, ty = (st / tl.nsx) * kSTY + w / kSTX;
    if (tx >= tl.ntx || ty >= tl.nty) return;   // a slot of a block past the edge
    const int X0 = rc.lo[0] + (int)tx * kTX, Y0 = rc.lo[1] + (int)ty * kTY;
    const int X1 = min(X0 + kTX, rc.hi[0]), Y1 = min(Y0 + kTY, rc.hi[1]);
    const int z0 = rc.lo[2] + (int)tz * 64 * kJ;
    const int z1 = min(z0 + 64 * kJ, rc.hi[2]);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int NT = (int)gc.NT;
    // this lane's setup slot: cell si of the wave, column ss of the cell
    const int si = lane % kC, ss = lane / kC;
    int cgi[kC], cgj[kC];
    bool cval[kC];
#pragma unroll
    for (int i = 0; i < kC; i++) {
        const int c = wave + kW * i;
        cgi[i] = X0 + c % kTX;
        cgj[i] = Y0 + c / kTX;
        cval[i] = cgi[i] < X1 && cgj[i] < Y1;
    }
    int s_ci = 0, s_cj = 0, s_cl = 0;
    bool s_cv = false;
#pragma unroll
    for (int i = 0; i < kC; i++)
        if (si == i) { s_ci = cgi[i]; s_cj = cgj[i]; s_cv = cval[i]; s_cl = wave + kW * i; }
    const int s_lx = s_cl % kTX, s_ly = s_cl / kTX;
    double acc[kC][kJ];
#pragma unroll
    for (int i = 0; i < kC; i++)
#pragma unroll
        for (int j = 0; j < kJ; j++) acc[i][j] = 0.0;
    double topacc = 0.0;   // G at z1 of this lane's setup cell (cell si)
    const float zf1 = (float)z1;

    // load item (view p, columns g0 .. g0 + ng - 1) into buffer B: 1/dz and P
    // rows in 1 KB pieces (16 bytes per lane, straight into LDS), the last piece
    // of each row set with only the lanes it needs
    auto stage = [&](uint32_t p, const GWin& wn, const GVu& g, int g0, int ng, GBuf& B) {
        // the cell corners' projections (column coordinate), written before the
        // loads into LDS are issued (no wait on them)
        if (tid < kNQc) {
            const int cx = tid % (kTX + 1), cy = tid / (kTX + 1);
            const float dx = (float)min(X0 + cx, rc.hi[0]) - g.sx;
            const float dy = (float)min(Y0 + cy, rc.hi[1]) - g.sy;
            B.qc[tid] = -(g.ax * dy - g.ay * dx) * frcp(g.bx * dy - g.by * dx);
        }
        const int Wp = (wn.W + 31) & ~31;
        const int bytes = Wp * 8;                  // per row set (a multiple of 256)
        const int npc = (bytes + 1023) / 1024;     // pieces per row set
        const int lastl = (bytes - (npc - 1) * 1024) / 16;   // lanes of the last piece
        // the lane's byte offsets, opaque to the compiler so that the 64-bit
        // staging addresses are formed here, not held in registers across items
        uint32_t l16 = (uint32_t)lane * 16, l4 = (uint32_t)lane * 4;
        asm volatile("" : "+v"(l16), "+v"(l4));
        char* dst0 = reinterpret_cast<char*>(B.win);
        const char* izp = reinterpret_cast<const char*>(izt + (uint64_t)p * gc.cs + wn.tlo) + l16;
        const char* colp = reinterpret_cast<const char*>(pref + ((uint64_t)p * gc.NQ + g0) * gc.cs + wn.tlo) + l16;
        // row set 0: 1/dz; row set c + 1: P of column g0 + c
        for (int e = wave; e < (ng + 1) * npc; e += kW) {
            const int c = e / npc, i = e - c * npc;
            const char* src = c == 0 ? izp + i * 1024 : colp + (uint64_t)(c - 1) * gc.cs * 8 + i * 1024;
            if (i + 1 < npc || lane < lastl)
                __builtin_amdgcn_global_load_lds(src, dst0 + c * bytes + i * 1024, 16, 0, 0);
        }
        // the group's column records (80 B each)
        for (int ch = wave; ch * 64 < ng * 20; ch += kW) {
            const char* src = reinterpret_cast<const char*>(gcols + (uint64_t)p * gc.NQ + g0) + ch * 256 + l4;
            __builtin_amdgcn_global_load_lds(src, reinterpret_cast<char*>(B.cols) + ch * 256, 4, 0, 0);
        }
    };

    // the setup and main passes of item (p, g0 .. g1) from buffer B
    auto work = [&](const GWin& wn, const GVu& g, int g0, int g1, const GBuf& B) {
        const int W = wn.W, tlo = wn.tlo;
        const int Wp = (W + 31) & ~31;
        // k - sz per slot, exact in FP64 (k an integer below 2^24, sz a float)
        double kzd[kJ];
#pragma unroll
        // k - sz per slot in FP64 with sz to FP64 precision (sz + szlo): a ray nearly
        // parallel to the planes crosses plane k at (k - sz) l / dz, so an FP32 sz
        // would move its crossing by up to 1e-5 l / dz voxels
        for (int jj = 0; jj < kJ; jj++) kzd[jj] = ((double)(z0 + lane + 64 * jj) - (double)g.sz) - (double)g.szlo;
        const double kz1d = ((double)z1 - (double)g.sz) - (double)g.szlo;
        const float kzspan = fmaxf(fabsf((float)z0 - g.sz), fabsf(zf1 - g.sz));
        const float Wm2 = (float)(W - 2);
        const float msz = -g.sz;
        // When every boundary of the tile lies at or above the source (z0 >= sz), a
        // ray that does not rise is entirely below each plane k, and the ray tf that
        // reaches plane k inside a cell rises: tf >= t0 (the first rising row), F is
        // the part before the crossing, and kappa_a - kappa_b = (k - sz)(alpha_a -
        // alpha_b) has one sign.  At or below the source (z1 - 1 <= sz) the same holds
        // for falling rays: tf < t0 and F is the part after the crossing.  Either
        // way one FMA gives kappa and the clamp of tf carries the side.  The window
        // must hold a row of that side.
        const int t0w = g.t0 - tlo;
        const bool upv = (float)z0 - g.sz >= 0.f && t0w <= W - 2;
        const bool dnv = !upv && (float)(z1 - 1) - g.sz <= 0.f && t0w >= 1;
        const float tlo_up = (float)max(0, t0w), thi_dn = (float)min(W - 2, t0w - 1);
        const char* Pb = reinterpret_cast<const char*>(B.win);   // 1/dz rows at 0, column c at (c + 1) Wp 8
        const double* Iw = B.win;
        // ---- setup pass: lane (si, ss) takes column qa(cell si) + ss ----
        float dh = 0.f, aa = 0.f, ba = 0.f, ab = 0.f, bbt = 0.f, rlq = 0.f, yhd = 0.f, ifh = 0.f;
        double hq = 0.0;
        int pb_ = 0, mp = 0, pcol = 0;
        bool valid = false;
        {
            const float c00 = B.qc[s_ly * (kTX + 1) + s_lx], c10 = B.qc[s_ly * (kTX + 1) + s_lx + 1];
            const float c01 = B.qc[(s_ly + 1) * (kTX + 1) + s_lx], c11 = B.qc[(s_ly + 1) * (kTX + 1) + s_lx + 1];
            const float cmin = fminf(fminf(c00, c10), fminf(c01, c11));
            const float cmax = fmaxf(fmaxf(c00, c10), fmaxf(c01, c11));
            const int qa = max(g0, (int)floorf(cmin - 1e-3f));
            const int qb = min(g1, (int)ceilf(cmax + 1e-3f));
            const int q = qa + ss;
            if (s_cv && q <= qb) {
                // the column record, all 80 bytes in one block of reads with one wait
                // (read field by field, its reads sit in the branches below, each with
                // its own wait)
                const uint32_t ca =
                    (uint32_t)(uintptr_t)(__attribute__((address_space(3))) const char*)(const char*)&B.cols[q - g0];
                float4 a0, a1, a2, a3, a4;
                asm volatile(
                    "ds_read_b128 %0, %5\n\t"
                    "ds_read_b128 %1, %5 offset:16\n\t"
                    "ds_read_b128 %2, %5 offset:32\n\t"
                    "ds_read_b128 %3, %5 offset:48\n\t"
                    "ds_read_b128 %4, %5 offset:64\n\t"
                    "s_waitcnt lgkmcnt(0)"
                    : "=&v"(a0), "=&v"(a1), "=&v"(a2), "=&v"(a3), "=&v"(a4)   // none shares the address register
                    : "v"(ca)
                    : "memory");
                // GCol: r[2] c[2] | P0[2] sg[2] | ifix[2] hs hd | l hc gam | yh ifixh lods pad
                const float Gr[2] = {a0.x, a0.y}, Gc[2] = {a0.z, a0.w}, GP0[2] = {a1.x, a1.y};
                const int Gsg[2] = {__float_as_int(a1.z), __float_as_int(a1.w)};
                const int Gif[2] = {__float_as_int(a2.x), __float_as_int(a2.y)};
                float h0 = a2.z, h1 = a2.w;
                const float Gl = a3.x, Ghc = a3.y;
                const double Ggam =
                    __longlong_as_double((long long)(((uint64_t)__float_as_uint(a3.w) << 32) | __float_as_uint(a3.z)));
                const float Gyh = a4.x, Glods = a4.z;
                const int Gifh = __float_as_int(a4.y);
                const int blo[2] = {s_ci, s_cj}, bhi[2] = {s_ci + 1, s_cj + 1};
#pragma unroll
                for (int a = 0; a < 2; a++) {
                    if (Gsg[a] == 0) {
                        if (Gif[a] != blo[a]) h1 = -AP_BIG;
                        continue;
                    }
                    const float ein = (float)(Gsg[a] > 0 ? blo[a] : bhi[a]) - GP0[a];
                    const float eout = (float)(Gsg[a] > 0 ? bhi[a] : blo[a]) - GP0[a];
                    h0 = fmaxf(h0, tcross(ein, Gr[a], Gc[a]));
                    h1 = fminf(h1, tcross(eout, Gr[a], Gc[a]));
                }
                if (h0 < h1) {
                    valid = true;
                    dh = h1 - h0;
                    aa = Glods * frcp(h0 + Ghc);
                    ab = Glods * frcp(h1 + Ghc);
                    ba = __builtin_fmaf(-g.sz, aa, wn.kc);
                    bbt = __builtin_fmaf(-g.sz, ab, wn.kc);
                    // the crossing in units of the chord: x' = (k - sz) / dz_t * l / dh - H,
                    // H = (hc / l) (l / dh) + h_a / dh.  A chord too short to divide by
                    // keeps x' = 0, an error of at most dh yhat_tf.
                    if (dh > 1e-20f) {
                        const double rd = 1.0 / (double)dh;
                        rlq = (float)((double)Gl * rd);
                        hq = fma(Ggam, (double)rlq, (double)h0 * rd);
                    }
                    pb_ = (q - g0 + 1) * Wp * 8;
                    pcol = q - g0;
                    yhd = Gyh * dh;
                    ifh = (float)Gifh;
                    // rays straddling one plane inside the cell lie in
                    // (min kappa, max kappa], |kappa_a - kappa_b| = |k - sz| |alpha_a - alpha_b|:
                    // more than one of them needs the general evaluation
                    mp = kzspan * fabsf(aa - ab) >= 0.99f ? 1 : 0;
                }
            }
        }
        // The part of ray t's chord below plane k, in units of the chord:
        // clamp((k - sz) / dz_t * l / dh - H, 0, 1) for a rising ray (one FP64 FMA with
        // the clamp modifier), one minus that for a falling one.
        auto below = [](double kz, double iz, float rl_, double h_) -> double {
            return __builtin_fmin(__builtin_fmax(fma(kz * iz, (double)rl_, -h_), 0.0), 1.0);
        };
        // G(k) of one (cell, column) pair at boundary kf (kz = kf - sz in FP64), added to a:
        // dh (P[tf] + F' (P[tf + 1] - P[tf]))
        auto gacc = [&](double& a, float kf, double kz, float dh_, float aa_, float ba_, float ab_, float bb_,
                        float rl_, double h_, int pbb) {
            const float ka = __builtin_fmaf(kf, aa_, ba_), kb = __builtin_fmaf(kf, ab_, bb_);
            const int tf = cvt_flr(__builtin_amdgcn_fmed3f(fminf(ka, kb), 0.f, Wm2));
            const double P0 = *reinterpret_cast<const double*>(Pb + pbb + tf * 8);
            const double P1 = *reinterpret_cast<const double*>(Pb + pbb + tf * 8 + 8);
            const double iz = Iw[tf];
            const double x = below(kz, iz, rl_, h_);
            const double F = iz > 0.0 ? x : 1.0 - x;
            a = fma((double)dh_, fma(F, P1 - P0, P0), a);
        };
        // the same with every ray that may straddle plane k
        auto gacc_multi = [&](double& a, float kf, double kz, float dh_, float aa_, float ba_, float ab_, float bb_,
                              float rl_, double h_, int pbb) {
            const float ka = __builtin_fmaf(kf, aa_, ba_), kb = __builtin_fmaf(kf, ab_, bb_);
            const int tf = cvt_flr(__builtin_amdgcn_fmed3f(fminf(ka, kb), 0.f, Wm2));
            const int te = cvt_flr(__builtin_amdgcn_fmed3f(fmaxf(ka, kb), (float)tf, Wm2));
            const double d = (double)dh_;
            a = fma(d, *reinterpret_cast<const double*>(Pb + pbb + tf * 8), a);
            for (int t = tf; t <= te; t++) {
                const double iz = Iw[t];
                const double x = below(kz, iz, rl_, h_);
                const double F = iz > 0.0 ? x : 1.0 - x;
                const double P0 = *reinterpret_cast<const double*>(Pb + pbb + t * 8);
                const double P1 = *reinterpret_cast<const double*>(Pb + pbb + t * 8 + 8);
                a = fma(d * F, P1 - P0, a);
            }
        };
        // G(k) on one side of the source (kap: kappa at the end that is the smaller
        // at every boundary of the tile); rising rays: dh (P[tf] + F' (P[tf+1] - P[tf])),
        // falling rays: dh (P[tf+1] - x' (P[tf+1] - P[tf])) with F' = 1 - x'
        auto gup = [&](double& a, float kap, double kz, double dh_, float rl_, double h_, int pbb) {
            const int tf = cvt_flr(__builtin_amdgcn_fmed3f(kap, tlo_up, Wm2));
            const double P0 = *reinterpret_cast<const double*>(Pb + pbb + tf * 8);
            const double P1 = *reinterpret_cast<const double*>(Pb + pbb + tf * 8 + 8);
            const double iz = Iw[tf];
            a = fma(dh_, fma(below(kz, iz, rl_, h_), P1 - P0, P0), a);
        };
        auto gdn = [&](double& a, float kap, double kz, double dh_, float rl_, double h_, int pbb) {
            const int tf = cvt_flr(__builtin_amdgcn_fmed3f(kap, 0.f, thi_dn));
            const double P0 = *reinterpret_cast<const double*>(Pb + pbb + tf * 8);
            const double P1 = *reinterpret_cast<const double*>(Pb + pbb + tf * 8 + 8);
            const double iz = Iw[tf];
            a = fma(dh_, fma(-below(kz, iz, rl_, h_), P1 - P0, P1), a);
        };
        {
            // flags: bit 24 several rays may straddle a plane, bit 25 the
            // column has a ray lying in a z plane, bits 26-29 its column in
            // the group; bits 0-23 the byte offset of its rows in the window
            float* R = prec(wave, lane);
            const uint64_t gbits = (uint64_t)__double_as_longlong(hq);
            const int flags = pb_ | (mp << 24) | ((yhd != 0.f ? 1 : 0) << 25) | (pcol << 26);
            reinterpret_cast<float2*>(R)[0] = make_float2(rlq, dh);
            reinterpret_cast<float2*>(R)[1] = make_float2(aa, ab);
            reinterpret_cast<float2*>(R)[2] = make_float2(0.f, __int_as_float(flags));
            reinterpret_cast<float2*>(R)[3] =
                make_float2(__uint_as_float((uint32_t)gbits), __uint_as_float((uint32_t)(gbits >> 32)));
        }
        if (valid) {
            double gt = 0.0;
            if (mp)
                gacc_multi(gt, zf1, kz1d, dh, aa, ba, ab, bbt, rlq, hq, pb_);
            else
                gacc(gt, zf1, kz1d, dh, aa, ba, ab, bbt, rlq, hq, pb_);
            if (yhd != 0.f && zf1 > ifh) gt += (double)yhd;
            topacc += gt;
        }
        // ---- main pass: every lane, every pair of cell i ----
#pragma unroll
        for (int i = 0; i < kC; i++) {
            uint64_t m = __ballot(valid && si == i);
            while (m) {
                const int src = __builtin_ctzll(m);
                m &= m - 1;
                // the pair's record, read by every lane at one address (a broadcast).
                // The read is written out, with its wait in the same block, so that
                // the compiler does not also wait for the next item's window loads
                // (its alias information does not reach every LDS read)
                const uint32_t ra = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) float*)prec(wave, src);
                float4 q0, q1;
                asm volatile(
                    "ds_read2_b64 %0, %2 offset1:1\n\t"
                    "ds_read2_b64 %1, %2 offset0:2 offset1:3\n\t"
                    "s_waitcnt lgkmcnt(0)"
                    : "=&v"(q0), "=&v"(q1)   // early clobber: no output shares the address register
                    : "v"(ra)
                    : "memory");
                const float rl = q0.x, rdh = q0.y, raa = q0.z, rab = q0.w;
                const double rdhd = (double)rdh;
                const int rfl = __float_as_int(q1.y);
                const int rpb = rfl & 0xFFFFFF;
                const double rh = __longlong_as_double(
                    (long long)(((uint64_t)__float_as_uint(q1.w) << 32) | __float_as_uint(q1.z)));
                if (rfl & (1 << 24)) {
                    const float rba = __builtin_fmaf(msz, raa, wn.kc), rbb = __builtin_fmaf(msz, rab, wn.kc);
                    for (int jj = 0; jj < kJ; jj++) {
                        const float kf = (float)(z0 + lane + 64 * jj);
                        gacc_multi(acc[i][jj], kf, kzd[jj], rdh, raa, rba, rab, rbb, rl, rh, rpb);
                    }
                } else if (upv) {
                    // kappa_a - kappa_b has the sign of alpha_a - alpha_b
                    const float ram = raa <= rab ? raa : rab;
                    const float rbm = __builtin_fmaf(msz, ram, wn.kc);
#pragma unroll
                    for (int jj = 0; jj < kJ; jj++)
                        gup(acc[i][jj], __builtin_fmaf((float)(z0 + lane + 64 * jj), ram, rbm), kzd[jj], rdhd, rl,
                            rh, rpb);
                } else if (dnv) {
                    // kappa_a - kappa_b has the sign of alpha_b - alpha_a
                    const float ram = raa >= rab ? raa : rab;
                    const float rbm = __builtin_fmaf(msz, ram, wn.kc);
#pragma unroll
                    for (int jj = 0; jj < kJ; jj++)
                        gdn(acc[i][jj], __builtin_fmaf((float)(z0 + lane + 64 * jj), ram, rbm), kzd[jj], rdhd, rl,
                            rh, rpb);
                } else {
                    const float rba = __builtin_fmaf(msz, raa, wn.kc), rbb = __builtin_fmaf(msz, rab, wn.kc);
#pragma unroll
                    for (int jj = 0; jj < kJ; jj++)
                        gacc(acc[i][jj], (float)(z0 + lane + 64 * jj), kzd[jj], rdh, raa, rba, rab, rbb, rl, rh, rpb);
                }
                if (rfl & (1 << 25)) {
                    // rays lying in a z plane: their voxel row is below the plane
                    const GCol& G = B.cols[(rfl >> 26) & 15];
                    const double ryh = (double)(G.yh * rdh);
                    const float rif = (float)G.ifixh;
#pragma unroll
                    for (int jj = 0; jj < kJ; jj++)
                        if ((float)(z0 + lane + 64 * jj) > rif) acc[i][jj] += ryh;
                }
            }
        }
    };

    // a window record read as wave-uniform values (scalar registers), so that the
    // view records read with them are scalar loads: a vector load in the work
    // phase would wait for the in-flight staging loads (one in-order counter)
    auto win_at = [&](int idx) -> GWin {
        const GWin w = Win[idx];
        GWin u;
        u.qlo = __builtin_amdgcn_readfirstlane(w.qlo);
        u.qhi = __builtin_amdgcn_readfirstlane(w.qhi);
        u.tlo = __builtin_amdgcn_readfirstlane(w.tlo);
        u.W = __builtin_amdgcn_readfirstlane(w.W);
        u.kc = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(w.kc)));
        u.ncg = __builtin_amdgcn_readfirstlane(w.ncg);
        u.pad[0] = u.pad[1] = 0;
        return u;
    };

    // view record idx of the batch, read as wave-uniform values
    auto vw_at = [&](int idx) -> GVu {
        const uint32_t va = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) float*)(Vw + idx * 8);
        float4 v0, v1;
        asm volatile(
            "ds_read2_b64 %0, %2 offset1:1\n\t"
            "ds_read2_b64 %1, %2 offset0:2 offset1:3\n\t"
            "s_waitcnt lgkmcnt(0)"
            : "=&v"(v0), "=&v"(v1)   // early clobber: no output shares the address register
            : "v"(va)
            : "memory");
        auto u = [](float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); };
        GVu g;
        g.sz = u(v0.x);
        g.sx = u(v0.y);
        g.sy = u(v0.z);
        g.t0 = __builtin_amdgcn_readfirstlane(__float_as_int(v0.w));
        g.ax = u(v1.x);
        g.ay = u(v1.y);
        g.bx = u(v1.z);
        g.by = u(v1.w);
        g.szlo = u(VwLo[idx]);
        return g;
    };

    for (uint32_t pb = 0; pb < gc.V; pb += 64) {
        // ---- windows and view records of views pb .. pb + 63, one per lane of wave 0 ----
        __syncthreads();   // the previous batch's readers of Win and Vw are done
        if (wave == 0) {
            GWin wn;
            wn.qlo = 0; wn.qhi = -1; wn.tlo = 0; wn.W = 2; wn.kc = 0.f; wn.ncg = 1; wn.pad[0] = wn.pad[1] = 0;
            const uint32_t p = pb + (uint32_t)lane;
            float4 v0 = make_float4(0.f, 0.f, 0.f, 0.f), v1 = v0;
            float szlo = 0.f;
            if (p < gc.V && gviews[p].ok) {
                const GView g = gviews[p];
                v0 = make_float4(g.sz, g.sx, g.sy, __int_as_float(g.t0));
                v1 = make_float4(g.ax, g.ay, g.bx, g.by);
                szlo = g.szlo;
                float qmin = AP_BIG, qmax = -AP_BIG, dmax = 0.f;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const float x = (float)((q & 1) ? X1 : X0), y = (float)((q & 2) ? Y1 : Y0);
                    const float dx = x - g.sx, dy = y - g.sy;
                    const float qc = -(g.ax * dy - g.ay * dx) * frcp(g.bx * dy - g.by * dx);
                    qmin = fminf(qmin, qc);
                    qmax = fmaxf(qmax, qc);
                    dmax = fmaxf(dmax, dx * dx + dy * dy);
                }
                const int qlo = max(0, (int)floorf(qmin - 1e-3f));
                const int qhi = min((int)gc.NQ - 1, (int)ceilf(qmax + 1e-3f));
                if (qmax >= -1.f && qmin <= (float)gc.NQ && qlo <= qhi) {
                    // horizontal distances: nearest point of the tile box, farthest corner
                    const float ex = fmaxf(fmaxf((float)X0 - g.sx, g.sx - (float)X1), 0.f);
                    const float ey = fmaxf(fmaxf((float)Y0 - g.sy, g.sy - (float)Y1), 0.f);
                    const float hmin = fmaxf(__builtin_amdgcn_sqrtf(ex * ex + ey * ey) * 0.999f, 1e-3f);
                    const float hmax = __builtin_amdgcn_sqrtf(dmax) * 1.001f;
                    // source-pixel distance over the window's columns
                    auto lcol = [&](float q) {
                        const float x = g.ax + q * g.bx, y = g.ay + q * g.by;
                        return __builtin_amdgcn_sqrtf(x * x + y * y);
                    };
                    const float bb = g.bx * g.bx + g.by * g.by;
                    const float qv = fminf(fmaxf(-(g.ax * g.bx + g.ay * g.by) * frcp(bb), (float)qlo), (float)qhi);
                    const float la = lcol((float)qlo), lb = lcol((float)qhi), lv = lcol(qv);
                    const float lmin = fminf(fminf(la, lb), lv) * 0.999f, lmax = fmaxf(la, lb) * 1.001f;
                    const float rmin = lmin * frcp(hmax), rmax = lmax * frcp(hmin);
                    const float kz0 = (float)z0 - g.sz, kz1 = zf1 - g.sz;
                    const float k00 = kz0 * rmin, k01 = kz0 * rmax, k10 = kz1 * rmin, k11 = kz1 * rmax;
                    const float ids = frcp(g.dzs);
                    const float kmn = (fminf(fminf(k00, k01), fminf(k10, k11)) - g.dz0) * ids;
                    const float kmx = (fmaxf(fmaxf(k00, k01), fmaxf(k10, k11)) - g.dz0) * ids;
                    const int tlo = (int)fminf(fmaxf(floorf(kmn) - 1.f, 0.f), (float)NT);
                    const int thi = (int)fminf(fmaxf(floorf(kmx) + 3.f, 0.f), (float)NT);
                    wn.qlo = qlo;
                    wn.qhi = qhi;
                    wn.tlo = tlo;
                    wn.W = thi - tlo + 2;   // P[tlo .. thi + 1]
                    wn.kc = 1.f - (float)tlo - g.dz0 / g.dzs;
                    // columns per group: the window holds Wp rows of 1/dz and of P per column
                    const int Wp = (wn.W + 31) & ~31;
                    wn.ncg = min(kGMaxCols, kGWin / (Wp * 8) - 1);
                }
            }
            Win[lane] = wn;
            reinterpret_cast<float4*>(Vw)[lane * 2] = v0;
            reinterpret_cast<float4*>(Vw)[lane * 2 + 1] = v1;
            VwLo[lane] = szlo;
        }
        __syncthreads();
        const uint32_t pe = min(gc.V, pb + 64);
        // the batch's items (view, column group) in order, the next one loading
        // while the current one is worked on
        uint32_t p = pb;
        GWin wn = win_at(0);
        while (p < pe && wn.qlo > wn.qhi) {   // the first view that meets the tile
            p++;
            if (p < pe) wn = win_at((int)(p - pb));
        }
        if (p >= pe) continue;
        int g0 = wn.qlo;
        GVu vw = vw_at((int)(p - pb));
        stage(p, wn, vw, g0, min(wn.qhi, g0 + wn.ncg - 1) - g0 + 1, Buf0);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        // one item: stage the next into Bn, work on this one from Bc.  Two named
        // buffers (not one indexed array) let the compiler see that the work's LDS
        // reads do not touch the loads in flight, so it does not wait for them.
        auto step = [&](GBuf& Bc, GBuf& Bn) {
            const int g1 = min(wn.qhi, g0 + wn.ncg - 1);
            // the next item: the view's next group, else the next view that meets the tile
            uint32_t pn = p;
            int gn = g1 + 1;
            GWin wnn = wn;
            if (gn > wn.qhi) {
                pn = p + 1;
                if (pn < pe) wnn = win_at((int)(pn - pb));
                while (pn < pe && wnn.qlo > wnn.qhi) {
                    pn++;
                    if (pn < pe) wnn = win_at((int)(pn - pb));
                }
                gn = wnn.qlo;
            }
            GVu vwn = vw;
            if (pn != p && pn < pe) vwn = vw_at((int)(pn - pb));
            if (pn < pe) stage(pn, wnn, vwn, gn, min(wnn.qhi, gn + wnn.ncg - 1) - gn + 1, Bn);
            work(wn, vw, g0, g1, Bc);
            __builtin_amdgcn_s_waitcnt(0);   // this wave's loads of the next item have landed
            __syncthreads();                 // all have, and the current buffer's readers are done
            p = pn;
            g0 = gn;
            wn = wnn;
            vw = vwn;
        };
        while (p < pe) {
            step(Buf0, Buf1);
            if (p >= pe) break;
            step(Buf1, Buf0);
        }
    }
    // ---- G at z1 per cell: sum over the lanes of the cell's setup slots ----
    for (int o = kC; o < 64; o <<= 1) topacc += __shfl_xor(topacc, o, 64);
    // ---- voxel values: differences of neighbouring boundaries ----
#pragma unroll
    for (int i = 0; i < kC; i++) {
        const double top = __shfl(topacc, i, 64);
        if (!cval[i]) continue;
#pragma unroll
        for (int jj = 0; jj < kJ; jj++) {
            const int k = z0 + lane + 64 * jj;
            double nxt = __shfl_down(acc[i][jj], 1, 64);
            if (jj + 1 < kJ) {
                const double n0 = __shfl(acc[i][jj + 1], 0, 64);
                if (lane == 63) nxt = n0;
            }
            if (k + 1 == z1) nxt = top;
            if (k < z1) {
                const float v = (float)((nxt - acc[i][jj]) * scale);
                const int64_t ov = (int64_t)cgi[i] * vsx + (int64_t)cgj[i] * vsy + (int64_t)k * vsz;
                const float mv = mult ? to_f32(mult[(int64_t)cgi[i] * msx + (int64_t)cgj[i] * msy + (int64_t)k * msz])
                                      : 1.f;
                vol[ov] = from_f32<Tv>(__builtin_fmaf(mv, v, to_f32(vol[ov])));
            }
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
uint64_t gather_workspace_bytes(uint32_t V, uint32_t nu, uint32_t nv, uint32_t ssu, uint32_t ssv) {
    if (V == 0 || nu == 0 || nv == 0) return 0;
    return (uint64_t)g_layout(V, nu * ssu, nv * ssv).total;
}

bool gather_supported(uint32_t nv, uint32_t ssv) {
    return (uint64_t)nv * ssv + 34 <= (uint64_t)kGRowMax;
}

__device__ bool gather_view_ok(const RayCfg& rc, const ap_proj_spec* __restrict__ specs, uint32_t p, uint32_t ssu,
                               uint32_t ssv, uint32_t nv) {
    GView gv;
    return gather_view(rc, specs, p, ssu, ssv, nv * ssv, gv);
}

namespace {

constexpr int kGW = 16, kGCv = 4, kGJ = 4, kGTX = 8;

}  // namespace

int32_t launch_gather_adjoint(const RayCfg& rc, uint32_t V, uint32_t nu, uint32_t nv, int64_t s0, int64_t s1,
                              int64_t s2, uint32_t ssu, uint32_t ssv, double scale, const ap_proj_spec* specs,
                              const ap_array3* vol_desc, void* vol, int32_t mdt, const int64_t* ms,
                              const void* mult, int32_t sino_dtype, const void* sino, void* workspace,
                              uint64_t ws_bytes, hipStream_t s, const uint32_t** left) {
    if (!gather_supported(nv, ssv)) return AP_ERR_UNSUPPORTED;
    GCfg gc;
    gc.V = V;
    gc.nu = nu;
    gc.nv = nv;
    gc.ssu = ssu;
    gc.ssv = ssv;
    gc.NQ = nu * ssu;
    gc.NT = nv * ssv;
    gc.s0 = s0;
    gc.s1 = s1;
    gc.s2 = s2;
    const GLayout L = g_layout(V, gc.NQ, gc.NT);
    gc.cs = L.cs;
    if (!workspace || ws_bytes < L.total || ((uintptr_t)workspace & 255)) return AP_ERR_UNSUPPORTED;
    char* ws = (char*)workspace;
    GView* gviews = (GView*)(ws + L.views);
    GCol* gcols = (GCol*)(ws + L.cols);
    double* pref = (double*)(ws + L.pref);
    double* izt = (double*)(ws + L.iz);
    const uint64_t nwave = (uint64_t)V * gc.NQ;
    if ((nwave + 3) / 4 >= (1ull << 31)) return AP_ERR_UNSUPPORTED;
    int32_t st = dispatch_dtype(sino_dtype, [&](auto tp) -> int32_t {
        using Tp = decltype(tp);
        if constexpr (std::is_same<Tp, double>::value || std::is_same<Tp, bf16_t>::value) {
            return ap_fail(AP_ERR_UNSUPPORTED, "sinogram dtype must be F32 or F16");
        } else {
            hipLaunchKernelGGL(k_gather_prep<Tp>, dim3((uint32_t)((nwave + 3) / 4)), dim3(256), 0, s, rc, gc, specs,
                               (const Tp*)sino, gviews, gcols, pref, izt);
            return ap_check_launch("gather_prep");
        }
    });
    if (st) return st;
    if (left) {
        uint32_t* cnt = (uint32_t*)(ws + L.count);
        hipLaunchKernelGGL(k_gather_count, dim3(1), dim3(256), 0, s, gviews, V, cnt);
        if ((st = ap_check_launch("gather_count"))) return st;
        *left = cnt;
    }
    constexpr int kTY = kGW * kGCv / kGTX;
    GTile tl;
    tl.ntx = (uint32_t)(rc.hi[0] - rc.lo[0] + kGTX - 1) / kGTX;
    tl.nty = (uint32_t)(rc.hi[1] - rc.lo[1] + kTY - 1) / kTY;
    tl.ntz = (uint32_t)(rc.hi[2] - rc.lo[2] + 64 * kGJ - 1) / (64 * kGJ);
    tl.nsx = (tl.ntx + kSTX - 1) / kSTX;
    tl.nsy = (tl.nty + kSTY - 1) / kSTY;
    const uint64_t total = (uint64_t)tl.nsx * tl.nsy * kST * tl.ntz;
    if (total >= (1ull << 31)) return ap_fail(AP_ERR_INVALID, "too many tiles");
    tl.total = (uint32_t)total;
    tl.per_xcd = (tl.total + kXcdG - 1) / kXcdG;
    dim3 grid(tl.per_xcd * kXcdG);
    return dispatch_dtype(vol_desc->dtype, [&](auto tv) -> int32_t {
        using Tv = decltype(tv);
        return dispatch_dtype(mdt, [&](auto tm) -> int32_t {
            using Tm = decltype(tm);
            if constexpr (std::is_same<Tv, double>::value || std::is_same<Tm, double>::value) {
                return ap_fail(AP_ERR_UNSUPPORTED, "unsupported dtype combination");
            } else {
                hipLaunchKernelGGL((k_gather_adjoint<Tv, Tm, kGW, kGCv, kGJ, kGTX>), grid, dim3(kGW * 64), 0, s, rc,
                                   gc, tl, gviews, gcols, pref, izt, scale, (Tv*)vol, vol_desc->strides[0],
                                   vol_desc->strides[1], vol_desc->strides[2], (const Tm*)mult, ms[0], ms[1], ms[2]);
                return ap_check_launch("gather_adjoint");
            }
        });
    });
}

// ---------------------------------------------------------------------------
// allocation-free entry point
// ---------------------------------------------------------------------------
namespace {

// The gather's workspace for callers that cannot allocate: 64 MiB per GPU,
// held from the first call on that GPU until the process ends.  Host code
// must not pass the symbol itself to a kernel or a copy (it names a host-side
// stand-in); the device address of each GPU's copy comes from
// hipGetSymbolAddress with that GPU current.
constexpr uint64_t kNoAllocBytes = 64ull << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int32_t scratch_address(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return ap_fail(AP_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return ap_fail(AP_ERR_UNSUPPORTED, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return ap_fail(AP_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return AP_OK;
}

}  // namespace

extern "C" int32_t ap_backproject_matched_noalloc(const ap_config_cb_project* cfg, const ap_proj_spec* specs,
                                                  uint32_t num_projs, const ap_vol_tex_chunk_spec* chunk,
                                                  const ap_array3* vol_desc, void* vol, const ap_array3* mult_desc,
                                                  const void* multiplier, const ap_array3* sino_desc,
                                                  const void* sino, float alpha, void* stream) {
    AP_REQUIRE(cfg && sino_desc, "NULL config or sinogram descriptor");
    AP_REQUIRE(cfg->ssu >= 1 && cfg->ssv >= 1, "ssu and ssv must be >= 1");
    const uint32_t nu = sino_desc->shape[1], nv = sino_desc->shape[2];
    if (gather_supported(nv, cfg->ssv) && num_projs > 0 && nu > 0 && nv > 0) {
        const uint64_t one = gather_workspace_bytes(1, nu, nv, cfg->ssu, cfg->ssv);
        if (one > kNoAllocBytes)
            return ap_fail(AP_ERR_UNSUPPORTED,
                           "one view of this detector (%u x %u, ss %u x %u) needs %llu bytes of scratch; the "
                           "allocation-free adjoint holds %llu (use ap_backproject_matched with a workspace)",
                           nu, nv, cfg->ssu, cfg->ssv, (unsigned long long)one,
                           (unsigned long long)kNoAllocBytes);
    }
    void* ws = nullptr;
    int32_t st = scratch_address(&ws);
    if (st) return st;
    return ap_backproject_matched(cfg, specs, num_projs, chunk, vol_desc, vol, mult_desc, multiplier, sino_desc,
                                  sino, alpha, ws, kNoAllocBytes, stream);
}
