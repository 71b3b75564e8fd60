// project.hip -- ray-driven Siddon forward projector (ap_project_voxel) and
// its exact transpose (ap_backproject_matched).
//
// Forward: one thread per detector pixel; a wavefront holds 64 consecutive
// pixels along v (the fastest sinogram axis, and z, the fastest volume axis,
// for the circular geometries this path serves), so the per-step gathers of a
// wave fall in a few cache lines and the sinogram stores are coalesced.  Each
// thread walks kBatch voxels before issuing their kBatch loads, so several
// gathers per wave are in flight.  Tiles are ordered view-group / u-tile /
// view / v-tile and split into one contiguous range per XCD, so the blocks
// resident at any time sweep the same slab of the volume for several
// neighbouring views out of the same L2.
//
// Adjoint: one workgroup per volume brick accumulates A^T y in LDS over all
// views.  Per view it visits the detector pixels whose rays can meet the
// brick (bounding box of the projected brick corners), clips each ray to the
// brick with the same crossing parameters as the forward walk and scatters
// weight*y into 64-bit fixed-point LDS sums (a first pass over the footprints
// finds max|y| for the scale); the brick is written to HBM once.  No global
// atomics, one write per voxel, and the result does not depend on the order
// of the adds.
#include "gather.h"

namespace {

constexpr int kWaveV = 64;   // pixels along v per wave
// Launch-shape variants (AP_TUNE_PROJECT_RAY): waves (pixel rows along u)
// per block and voxels walked per batch of gathers.
constexpr int kRayRowsU[3] = {4, 2, 1};
constexpr int kRayBatch[3] = {4, 4, 8};
constexpr int kViewGroup = 8;
constexpr int kXcd = 8;      // MI355X: 8 XCDs, workgroups dispatched round-robin

struct SinoCfg {
    uint32_t V, nu, nv;
    int64_t s0, s1, s2;      // element strides
    int64_t b0, b1, b2;      // element strides of the fused residual's b
};

struct TileMap {
    uint32_t nk, nj;         // v-tiles, u-tiles per view
    uint32_t total;          // padded tile count (multiple of view groups)
    uint32_t per_xcd;        // tiles per XCD
};

__device__ __forceinline__ bool map_tile(const TileMap& tm, uint32_t V, uint32_t L, uint32_t& p, uint32_t& jt,
                                         uint32_t& kt) {
    const uint32_t logical = (L % kXcd) * tm.per_xcd + L / kXcd;
    if (logical >= tm.total) return false;
    kt = logical % tm.nk;
    uint32_t r = logical / tm.nk;
    const uint32_t vi = r % kViewGroup;
    r /= kViewGroup;
    jt = r % tm.nj;
    const uint32_t vg = r / tm.nj;
    p = vg * kViewGroup + vi;
    return p < V;
}

__device__ __forceinline__ float sub_offset(uint32_t s, uint32_t ss) {
    return -0.5f + ((float)s + 0.5f) / (float)ss;
}

// kOne: one ray per pixel (ssu = ssv = 1), so the view's ray frame is dead
// during the walk (fewer registers live across it).
template <int kRowsU, int kBatch, typename Tv, typename Tp, typename IdxT, bool kOne>
__global__ __launch_bounds__(kWaveV* kRowsU) void k_project_voxel(
    RayCfg rc, SinoCfg sc, TileMap tm, float mba, uint32_t ssu, uint32_t ssv,
    const ap_proj_spec* __restrict__ specs, const Tv* __restrict__ vol, IdxT vsx, IdxT vsy, IdxT vsz,
    Tp* __restrict__ sino, const Tp* __restrict__ b, int skip_upright) {
    // A bounded grid strides over the tiles (a multiple of kXcd blocks, so a
    // block stays in its XCD's range).  The 64 lanes test the block's next 64
    // tiles at once (in range, view not taken by the column kernel) and the
    // block walks the ballot of live ones: tiles of skipped views cost a
    // 64th of one spec load each, not a workgroup launch or a serial load.
    // The test depends on threadIdx.x only, so every wave sees the same mask.
    const uint32_t lim = tm.per_xcd * kXcd;
    if constexpr (kOne) ssu = ssv = 1;
    for (uint32_t L0 = blockIdx.x; L0 < lim; L0 += 64u * gridDim.x) {
        bool live;
        {
            const uint32_t Lt = L0 + threadIdx.x * gridDim.x;
            uint32_t p, jt, kt;
            live = Lt < lim && map_tile(tm, sc.V, Lt, p, jt, kt) && !(skip_upright && view_is_upright(specs, p));
        }
        uint64_t mask = __ballot(live);
        while (mask) {
            const uint32_t L = L0 + (uint32_t)__builtin_ctzll(mask) * gridDim.x;
            mask &= mask - 1;
            uint32_t p, jt, kt;
            map_tile(tm, sc.V, L, p, jt, kt);
            const uint32_t k = kt * kWaveV + threadIdx.x;
            const uint32_t j = jt * kRowsU + threadIdx.y;
            if (k >= sc.nv || j >= sc.nu) continue;
            const ViewRays vr = make_view_rays(rc, specs, p);
            float tot = 0.f;
            for (uint32_t su = 0; su < ssu; su++) {
                float accu = 0.f;
                for (uint32_t sv = 0; sv < ssv; sv++) {
                    const float ju = (float)j + sub_offset(su, ssu);
                    const float kv = (float)k + sub_offset(sv, ssv);
                    Ray R;
                    if (!ray_setup(vr, rc.lo, rc.hi, ju, kv, R)) continue;
                    ray_start(R, R.t_enter, rc.lo, rc.hi);
                    Walker<IdxT> W;
                    W.init(R, R.t_enter, R.t_exit, vsx, vsy, vsz,
                           (IdxT)R.idx[0] * vsx + (IdxT)R.idx[1] * vsy + (IdxT)R.idx[2] * vsz);
                    float acc = 0.f;
                    for (int it = 0; it < rc.maxit; it += kBatch) {
                        float w[kBatch];
                        IdxT o[kBatch];
                        bool more = true;
        #pragma unroll
                        for (int q = 0; q < kBatch; q++) more = W.step(w[q], o[q]);
                        float v[kBatch];
        #pragma unroll
                        for (int q = 0; q < kBatch; q++) v[q] = to_f32(vol[o[q]]);
        #pragma unroll
                        for (int q = 0; q < kBatch; q++) acc = __builtin_fmaf(w[q], v[q], acc);
                        if (!more) break;
                    }
                    accu += acc;
                }
                tot += ssv > 1 ? accu / (float)ssv : accu;
            }
            if (ssu > 1) tot /= (float)ssu;
            tot *= (float)(1.0 / rc.inv_vw);   // voxel-length units -> physical
            const int64_t o = (int64_t)p * sc.s0 + (int64_t)j * sc.s1 + (int64_t)k * sc.s2;
            const float base =
                to_f32(b ? b[(int64_t)p * sc.b0 + (int64_t)j * sc.b1 + (int64_t)k * sc.b2] : sino[o]);
            sino[o] = from_f32<Tp>(__builtin_fmaf(mba, tot, base));
        }
    }
}

// ---------------------------------------------------------------- adjoint
constexpr int kBX = 16, kBY = 16, kBZ = 16;    // brick (voxels), z fastest in LDS
constexpr int kAdjThreads = 256;
constexpr int kViewBatchAdj = 32;

struct BrickView {
    ViewRays vr;
    int j0, j1, k0, k1;   // detector footprint (inclusive), empty when j1 < j0
};

// x rounded to the nearest integer (ties to even), as a 64-bit two's
// complement value, for |x| < 2^51: adding 1.5 * 2^52 in FP64 leaves
// 2^51 + rint(x) in the mantissa field, so the bit pattern minus that of
// 1.5 * 2^52 is rint(x).  (The low word of the constant is 0: one 32-bit
// subtraction on the high word.)
__device__ __forceinline__ unsigned long long fx_term64(float x) {
    const double d = (double)x + 6755399441055744.0;
    return (unsigned long long)__double_as_longlong(d) - 0x4338000000000000ull;
}

// Pixel coordinates of the central projection of X (voxel units rel. centre):
// X = S + lam (ju U + kv V + Q00 - S) solved with the dual basis of (U, V).
__device__ __forceinline__ bool project_point(const ViewRays& vd, const float X[3], float& ju, float& kv) {
    struct { float S[3], Q00[3], U[3], V[3]; } vr;
    for (int a = 0; a < 3; a++) {
        vr.S[a] = (float)vd.S[a];
        vr.Q00[a] = (float)vd.Q00[a];
        vr.U[a] = (float)vd.U[a];
        vr.V[a] = (float)vd.V[a];
    }
    const float n[3] = {vr.U[1] * vr.V[2] - vr.U[2] * vr.V[1], vr.U[2] * vr.V[0] - vr.U[0] * vr.V[2],
                        vr.U[0] * vr.V[1] - vr.U[1] * vr.V[0]};
    float xs[3], qs[3];
    for (int a = 0; a < 3; a++) {
        xs[a] = X[a] - vr.S[a];
        qs[a] = vr.Q00[a] - vr.S[a];
    }
    const float den = xs[0] * n[0] + xs[1] * n[1] + xs[2] * n[2];
    const float num = qs[0] * n[0] + qs[1] * n[1] + qs[2] * n[2];
    if (!(den * num > 0.f)) return false;    // point not in front of the source
    const float lam = num / den;
    float P[3];
    for (int a = 0; a < 3; a++) P[a] = vr.S[a] + lam * xs[a] - vr.Q00[a];
    const float nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    // ju = ((P x V) . n) / |n|^2,  kv = ((U x P) . n) / |n|^2
    const float pv[3] = {P[1] * vr.V[2] - P[2] * vr.V[1], P[2] * vr.V[0] - P[0] * vr.V[2],
                         P[0] * vr.V[1] - P[1] * vr.V[0]};
    const float up[3] = {vr.U[1] * P[2] - vr.U[2] * P[1], vr.U[2] * P[0] - vr.U[0] * P[2],
                         vr.U[0] * P[1] - vr.U[1] * P[0]};
    ju = (pv[0] * n[0] + pv[1] * n[1] + pv[2] * n[2]) / nn;
    kv = (up[0] * n[0] + up[1] * n[1] + up[2] * n[2]) / nn;
    return true;
}

template <typename Tv, typename Tp, typename Tm, typename IdxT>
// 4 waves per SIMD (at most 128 VGPRs): the 32 KB brick leaves room for 4 workgroups per CU
__global__ __launch_bounds__(kAdjThreads, 4) void k_backproject_matched(
    RayCfg rc, SinoCfg sc, float scale, uint32_t ssu, uint32_t ssv,
    const ap_proj_spec* __restrict__ specs, Tv* __restrict__ vol, IdxT vsx, IdxT vsy, IdxT vsz,
    const Tm* __restrict__ mult, IdxT msx, IdxT msy, IdxT msz, const Tp* __restrict__ sino,
    uint32_t nbx, uint32_t nby, uint32_t nbz, uint32_t per_xcd, int skip_gather, const uint32_t* gather_left) {
    // 64-bit fixed point: terms rint(w * y * scale * S), each an exact
    // rounding of the FP32 product to an integer.  T bounds the terms that
    // can reach one voxel (the sub-rays through the brick's footprints over
    // all views: a ray crosses a voxel once) and |w| <= sqrt(3) < 2 voxels,
    // so S = 2^61 / (T max|y * scale|) keeps every sum below 2^62 while the
    // terms keep their FP32 precision relative to their own size.  Integer
    // sums do not depend on the order of the adds: the result is
    // deterministic.
    __shared__ unsigned long long acc[kBX * kBY * kBZ];
    __shared__ BrickView bv[kViewBatchAdj];
    __shared__ float wred[kAdjThreads / 64];
    __shared__ double tred[kAdjThreads / 64];
    // every view went to the gather kernel: nothing left here (uniform)
    if (gather_left && *gather_left == 0u) return;
    const uint32_t L = blockIdx.x;
    const uint32_t logical = (L % kXcd) * per_xcd + L / kXcd;
    const uint32_t nbr = nbx * nby * nbz;
    if (logical >= nbr) return;
    // bricks ordered z fastest so an XCD's bricks form compact columns
    const uint32_t bk = logical % nbz, bj = (logical / nbz) % nby, bi = logical / (nbz * nby);
    int blo[3] = {rc.lo[0] + (int)bi * kBX, rc.lo[1] + (int)bj * kBY, rc.lo[2] + (int)bk * kBZ};
    int bhi[3] = {min(blo[0] + kBX, rc.hi[0]), min(blo[1] + kBY, rc.hi[1]), min(blo[2] + kBZ, rc.hi[2])};
    const int tid = threadIdx.x;
    for (int e = tid; e < kBX * kBY * kBZ; e += kAdjThreads) acc[e] = 0ull;
    const int maxit = kBX + kBY + kBZ + 3;
    const float half_u = 0.5f, half_v = 0.5f;
    // footprints of views [p0, p0 + nb) into bv (one thread per view)
    auto fill_views = [&](uint32_t p0, uint32_t nb) {
        GView gv;
        if (tid < (int)nb && skip_gather && gather_view(rc, specs, p0 + tid, ssu, ssv, sc.nv * ssv, gv)) {
            bv[tid].j0 = 0; bv[tid].j1 = -1; bv[tid].k0 = 0; bv[tid].k1 = -1;
        } else if (tid < (int)nb) {
            BrickView& B = bv[tid];
            B.vr = make_view_rays(rc, specs, p0 + tid);
            // footprint: bounding box of the projected brick corners, widened
            // by half a pixel for sub-pixel rays
            float jmin = AP_BIG, jmax = -AP_BIG, kmin = AP_BIG, kmax = -AP_BIG;
            bool ok = true;
            for (int q = 0; q < 8; q++) {
                const float X[3] = {(float)((q & 4) ? bhi[0] : blo[0]) - 0.5f * (float)(rc.lo[0] + rc.hi[0]),
                                    (float)((q & 2) ? bhi[1] : blo[1]) - 0.5f * (float)(rc.lo[1] + rc.hi[1]),
                                    (float)((q & 1) ? bhi[2] : blo[2]) - 0.5f * (float)(rc.lo[2] + rc.hi[2])};
                float ju, kv;
                if (!project_point(B.vr, X, ju, kv)) { ok = false; break; }
                jmin = fminf(jmin, ju); jmax = fmaxf(jmax, ju);
                kmin = fminf(kmin, kv); kmax = fmaxf(kmax, kv);
            }
            if (ok) {
                // 1e-3 pixel of slack absorbs the FP32 rounding of the projection
                B.j0 = max(0, (int)floorf(jmin - half_u - 1e-3f));
                B.j1 = min((int)sc.nu - 1, (int)ceilf(jmax + half_u + 1e-3f));
                B.k0 = max(0, (int)floorf(kmin - half_v - 1e-3f));
                B.k1 = min((int)sc.nv - 1, (int)ceilf(kmax + half_v + 1e-3f));
            } else {   // brick not entirely in front of the source: every pixel
                B.j0 = 0; B.j1 = (int)sc.nu - 1; B.k0 = 0; B.k1 = (int)sc.nv - 1;
            }
        }
    };
    // pass 1: largest |y| over the footprints sets the fixed-point scale
    float ymax = 0.f;
    double nterm = 0.0;   // footprint sub-rays of this thread's views
    for (uint32_t p0 = 0; p0 < sc.V; p0 += kViewBatchAdj) {
        const uint32_t nb = min((uint32_t)kViewBatchAdj, sc.V - p0);
        __syncthreads();
        fill_views(p0, nb);
        __syncthreads();
        if (tid < (int)nb && bv[tid].j1 >= bv[tid].j0 && bv[tid].k1 >= bv[tid].k0)
            nterm += (double)(bv[tid].j1 - bv[tid].j0 + 1) * (double)(bv[tid].k1 - bv[tid].k0 + 1);
        for (uint32_t q = 0; q < nb; q++) {
            const BrickView& B = bv[q];
            if (B.j1 < B.j0 || B.k1 < B.k0) continue;
            const int nk = B.k1 - B.k0 + 1;
            const int npx = (B.j1 - B.j0 + 1) * nk;
            const uint32_t p = p0 + q;
            for (int pi = tid; pi < npx; pi += kAdjThreads) {
                const int j = B.j0 + pi / nk, k = B.k0 + pi % nk;
                const float y = to_f32(sino[(int64_t)p * sc.s0 + (int64_t)j * sc.s1 + (int64_t)k * sc.s2]);
                ymax = fmaxf(ymax, fabsf(y));
                if (y != y) ymax = __builtin_inff();
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        ymax = fmaxf(ymax, __shfl_xor(ymax, o, 64));
        nterm += __shfl_xor(nterm, o, 64);
    }
    if ((tid & 63) == 0) {
        wred[tid >> 6] = ymax;
        tred[tid >> 6] = nterm;
    }
    __syncthreads();
    ymax = 0.f;
    nterm = 0.0;
    for (int w = 0; w < kAdjThreads / 64; w++) {
        ymax = fmaxf(ymax, wred[w]);
        nterm += tred[w];
    }
    nterm *= (double)ssu * (double)ssv;
    const float amax = ymax * fabsf(scale);
    if (amax == 0.f) return;   // block-uniform: nothing to add
    if (!(amax < __builtin_inff())) {
        // non-finite data: the sum is not representable; propagate NaN
        for (int e = tid; e < kBX * kBY * kBZ; e += kAdjThreads) {
            const int li = e / (kBY * kBZ), lj = (e / kBZ) % kBY, lk = e % kBZ;
            const int gi = blo[0] + li, gj = blo[1] + lj, gk = blo[2] + lk;
            if (gi >= bhi[0] || gj >= bhi[1] || gk >= bhi[2]) continue;
            vol[(IdxT)gi * vsx + (IdxT)gj * vsy + (IdxT)gk * vsz] = from_f32<Tv>(__builtin_nanf(""));
        }
        return;
    }
    // terms below 2^50 (the exact FP64 rounding of fx_term64), sums below 2^62
    const double S = fmin(2305843009213693952.0 / fmax(nterm, 1.0), 1125899906842624.0) / (double)amax;
    const double scaleS = (double)scale * S;
    for (uint32_t p0 = 0; p0 < sc.V; p0 += kViewBatchAdj) {
        const uint32_t nb = min((uint32_t)kViewBatchAdj, sc.V - p0);
        __syncthreads();
        fill_views(p0, nb);
        __syncthreads();
        for (uint32_t q = 0; q < nb; q++) {
            const BrickView& B = bv[q];
            if (B.j1 < B.j0 || B.k1 < B.k0) continue;
            const int nk = B.k1 - B.k0 + 1;
            const int npx = (B.j1 - B.j0 + 1) * nk;
            const uint32_t p = p0 + q;
            for (int pi = tid; pi < npx; pi += kAdjThreads) {
                const int j = B.j0 + pi / nk, k = B.k0 + pi % nk;
                const float y = to_f32(sino[(int64_t)p * sc.s0 + (int64_t)j * sc.s1 + (int64_t)k * sc.s2]);
                if (y == 0.f) continue;
                const float ys = (float)((double)y * scaleS);
                for (uint32_t su = 0; su < ssu; su++)
                    for (uint32_t sv = 0; sv < ssv; sv++) {
                        Ray R;
                        if (!ray_setup(B.vr, rc.lo, rc.hi, (float)j + sub_offset(su, ssu),
                                       (float)k + sub_offset(sv, ssv), R))
                            continue;
                        float tb0, tb1;
                        ray_box(R, blo, bhi, tb0, tb1);
                        if (!(tb0 < tb1)) continue;
                        ray_start(R, tb0, blo, bhi);
                        Walker<int> W;
                        W.init(R, tb0, tb1, kBY * kBZ, kBZ, 1,
                               ((R.idx[0] - blo[0]) * kBY + (R.idx[1] - blo[1])) * kBZ + (R.idx[2] - blo[2]));
                        for (int it = 0; it < maxit; it++) {
                            float w;
                            int o;
                            const bool more = W.step(w, o);
                            atomicAdd(&acc[o], fx_term64(w * ys));
                            if (!more) break;
                        }
                    }
            }
        }
    }
    __syncthreads();
    const double inv_s = 1.0 / (double)S;
    for (int e = tid; e < kBX * kBY * kBZ; e += kAdjThreads) {
        const int li = e / (kBY * kBZ), lj = (e / kBZ) % kBY, lk = e % kBZ;
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

int32_t launch_project_columns(const RayCfg& rc, uint32_t V, uint32_t nu, uint32_t nv, int64_t s0, int64_t s1,
                               int64_t s2, float mba, uint32_t ssu, uint32_t ssv, const ap_proj_spec* specs,
                               const ap_array3* vol_desc, const void* vol, int32_t sino_dtype, void* sino,
                               const void* b, const int64_t* bs, bool idx64, hipStream_t s);
namespace {

int32_t check_sino(const ap_array3* sd, const void* sino, uint32_t num_projs, SinoCfg& sc,
                   bool allow_f64) {
    int32_t st = ap_check_array3(sd, sino, "sino", true, false, allow_f64);
    if (st) return st;
    if (sd->shape[0] != num_projs)
        return ap_fail(AP_ERR_INVALID, "sino: %u views in the array but num_projs=%u",
                       sd->shape[0], num_projs);
    sc.V = sd->shape[0];
    sc.nu = sd->shape[1];
    sc.nv = sd->shape[2];
    sc.s0 = sc.b0 = sd->strides[0];
    sc.s1 = sc.b1 = sd->strides[1];
    sc.s2 = sc.b2 = sd->strides[2];
    return AP_OK;
}

}  // namespace

int32_t make_ray_cfg(const ap_config_cb_project* cfg, const ap_vol_tex_chunk_spec* chunk,
                     const ap_array3* vol_desc, RayCfg& rc) {
    AP_REQUIRE(cfg && chunk && vol_desc, "NULL config, chunk or volume descriptor");
    AP_REQUIRE(cfg->ssu >= 1 && cfg->ssv >= 1, "ssu and ssv must be >= 1");
    AP_REQUIRE(chunk->voxwid > 0.f, "voxwid must be positive");
    const uint32_t off[3] = {chunk->offsets.x, chunk->offsets.y, chunk->offsets.z};
    const uint32_t shp[3] = {chunk->shape.x, chunk->shape.y, chunk->shape.z};
    int ext = 0;
    for (int a = 0; a < 3; a++) {
        AP_REQUIRE((uint64_t)off[a] + shp[a] <= vol_desc->shape[a],
                   "chunk exceeds the volume along axis %d", a);
        AP_REQUIRE(shp[a] >= 1, "empty chunk along axis %d", a);
        AP_REQUIRE((uint64_t)off[a] + shp[a] < (1u << 24), "volume extent above 2^24 voxels");
        rc.lo[a] = (int)off[a];
        rc.hi[a] = (int)(off[a] + shp[a]);
        ext += (int)shp[a];
    }
    rc.tp[0] = chunk->texpos.x;
    rc.tp[1] = chunk->texpos.y;
    rc.tp[2] = chunk->texpos.z;
    rc.svf = cfg->shrink_volume_factor;
    rc.inv_vw = 1.0 / (double)chunk->voxwid;
    rc.maxit = ext + 3;
    return AP_OK;
}

extern "C" int32_t ap_project_voxel(const ap_config_cb_project* cfg, const ap_proj_spec* specs,
                                    uint32_t num_projs, const ap_vol_tex_chunk_spec* chunk,
                                    const ap_array3* vol_desc, const void* vol,
                                    const ap_array3* sino_desc, void* sino,
                                    const ap_array3* b_desc, const void* b, void* stream) {
    AP_REQUIRE(specs, "specs is NULL");
    int32_t st = ap_check_array3(vol_desc, vol, "volume", true, true, false);
    if (st) return st;
    SinoCfg sc;
    if ((st = check_sino(sino_desc, sino, num_projs, sc, false))) return st;
    if (b) {
        // the fused residual's b: same shape and dtype as sino, its own strides
        AP_REQUIRE(b_desc, "b is given without its descriptor");
        if ((st = ap_check_array3(b_desc, b, "b", true, false, false))) return st;
        for (int a = 0; a < 3; a++)
            AP_REQUIRE(b_desc->shape[a] == sino_desc->shape[a], "b shape != sino shape along axis %d", a);
        AP_REQUIRE(b_desc->dtype == sino_desc->dtype, "b dtype != sino dtype");
        AP_REQUIRE(b != sino || (b_desc->strides[0] == sino_desc->strides[0] &&
                                 b_desc->strides[1] == sino_desc->strides[1] &&
                                 b_desc->strides[2] == sino_desc->strides[2]),
                   "b aliases sino with other strides");
        sc.b0 = b_desc->strides[0];
        sc.b1 = b_desc->strides[1];
        sc.b2 = b_desc->strides[2];
    }
    RayCfg rc;
    if ((st = make_ray_cfg(cfg, chunk, vol_desc, rc))) return st;
    if (num_projs == 0 || sc.nu == 0 || sc.nv == 0) return AP_OK;
    const int64_t maxoff = array3_max_offset(vol_desc);
    const bool idx64 = maxoff >= (int64_t)INT32_MAX;
    const int var = ap_tune_variant(AP_TUNE_PROJECT_RAY);
    const uint32_t rows_u = (uint32_t)kRayRowsU[var];
    TileMap tm;
    tm.nk = (sc.nv + kWaveV - 1) / kWaveV;
    tm.nj = (sc.nu + rows_u - 1) / rows_u;
    const uint64_t groups = (num_projs + kViewGroup - 1) / kViewGroup;
    const uint64_t total = groups * kViewGroup * tm.nj * tm.nk;
    AP_REQUIRE(total < (1ull << 31), "too many tiles in one call");
    tm.total = (uint32_t)total;
    tm.per_xcd = (tm.total + kXcd - 1) / kXcd;
    dim3 block(kWaveV, rows_u, 1);
    dim3 grid(min(tm.per_xcd, 512u) * kXcd, 1, 1);
    hipStream_t s = (hipStream_t)stream;
    const float mba = cfg->multiply_before_adding;
    // upright views go to the column-coherent kernel; the general kernel
    // then only takes the remaining views
    const int64_t bs[3] = {sc.b0, sc.b1, sc.b2};
    st = launch_project_columns(rc, sc.V, sc.nu, sc.nv, sc.s0, sc.s1, sc.s2, mba, cfg->ssu, cfg->ssv, specs,
                                vol_desc, vol, sino_desc->dtype, sino, b, bs, idx64, s);
    if (st != AP_OK && st != AP_ERR_UNSUPPORTED) return st;
    const int skip = st == AP_OK ? 1 : 0;
    return dispatch_dtype(vol_desc->dtype, [&](auto tv) -> int32_t {
        using Tv = decltype(tv);
        if constexpr (std::is_same<Tv, double>::value) {
            return ap_fail(AP_ERR_UNSUPPORTED, "float64 volumes are not supported");
        } else {
            return dispatch_dtype(sino_desc->dtype, [&](auto tp) -> int32_t {
                using Tp = decltype(tp);
                if constexpr (std::is_same<Tp, double>::value || std::is_same<Tp, bf16_t>::value) {
                    return ap_fail(AP_ERR_UNSUPPORTED, "sinogram dtype must be F32 or F16");
                } else {
                    auto launch = [&](auto kern, auto idx) {
                        using IdxT = decltype(idx);
                        hipLaunchKernelGGL(kern, grid, block, 0, s, rc, sc, tm, mba, cfg->ssu, cfg->ssv, specs,
                                           (const Tv*)vol, (IdxT)vol_desc->strides[0], (IdxT)vol_desc->strides[1],
                                           (IdxT)vol_desc->strides[2], (Tp*)sino, (const Tp*)b, skip);
                    };
                    if (idx64) {
                        if (var == 0) launch(k_project_voxel<kRayRowsU[0], kRayBatch[0], Tv, Tp, int64_t, false>, int64_t{});
                        if (var == 1) launch(k_project_voxel<kRayRowsU[1], kRayBatch[1], Tv, Tp, int64_t, false>, int64_t{});
                        if (var == 2) launch(k_project_voxel<kRayRowsU[2], kRayBatch[2], Tv, Tp, int64_t, false>, int64_t{});
                    } else if (cfg->ssu == 1 && cfg->ssv == 1) {
                        if (var == 0) launch(k_project_voxel<kRayRowsU[0], kRayBatch[0], Tv, Tp, int32_t, true>, int32_t{});
                        if (var == 1) launch(k_project_voxel<kRayRowsU[1], kRayBatch[1], Tv, Tp, int32_t, true>, int32_t{});
                        if (var == 2) launch(k_project_voxel<kRayRowsU[2], kRayBatch[2], Tv, Tp, int32_t, true>, int32_t{});
                    } else {
                        if (var == 0) launch(k_project_voxel<kRayRowsU[0], kRayBatch[0], Tv, Tp, int32_t, false>, int32_t{});
                        if (var == 1) launch(k_project_voxel<kRayRowsU[1], kRayBatch[1], Tv, Tp, int32_t, false>, int32_t{});
                        if (var == 2) launch(k_project_voxel<kRayRowsU[2], kRayBatch[2], Tv, Tp, int32_t, false>, int32_t{});
                    }
                    return ap_check_launch("project_voxel");
                }
            });
        }
    });
}

extern "C" int32_t ap_backproject_matched(const ap_config_cb_project* cfg,
                                          const ap_proj_spec* specs, uint32_t num_projs,
                                          const ap_vol_tex_chunk_spec* chunk,
                                          const ap_array3* vol_desc, void* vol,
                                          const ap_array3* mult_desc, const void* multiplier,
                                          const ap_array3* sino_desc, const void* sino,
                                          float alpha, void* workspace, uint64_t workspace_bytes,
                                          void* stream) {
    AP_REQUIRE(specs, "specs is NULL");
    int32_t st = ap_check_array3(vol_desc, vol, "volume", true, true, false);
    if (st) return st;
    if (multiplier) {
        if ((st = ap_check_array3(mult_desc, multiplier, "multiplier", true, true, false))) return st;
        for (int a = 0; a < 3; a++)
            AP_REQUIRE(mult_desc->shape[a] == vol_desc->shape[a], "multiplier shape != volume shape");
    }
    SinoCfg sc;
    if ((st = check_sino(sino_desc, sino, num_projs, sc, false))) return st;
    RayCfg rc;
    if ((st = make_ray_cfg(cfg, chunk, vol_desc, rc))) return st;
    if (num_projs == 0 || sc.nu == 0 || sc.nv == 0) return AP_OK;
    const int64_t maxoff = array3_max_offset(vol_desc);
    const int64_t maxoffm = multiplier ? array3_max_offset(mult_desc) : 0;
    const bool idx64 = maxoff >= (int64_t)INT32_MAX || maxoffm >= (int64_t)INT32_MAX;
    const float scale = (float)((double)alpha * cfg->multiply_before_adding /
                                ((double)cfg->ssu * cfg->ssv) * chunk->voxwid);
    const uint32_t nbx = (chunk->shape.x + kBX - 1) / kBX;
    const uint32_t nby = (chunk->shape.y + kBY - 1) / kBY;
    const uint32_t nbz = (chunk->shape.z + kBZ - 1) / kBZ;
    const uint64_t nbr = (uint64_t)nbx * nby * nbz;
    AP_REQUIRE(nbr < (1ull << 31), "too many bricks");
    const uint32_t per_xcd = (uint32_t)((nbr + kXcd - 1) / kXcd);
    dim3 grid(per_xcd * kXcd, 1, 1);
    hipStream_t s = (hipStream_t)stream;
    const int32_t mdt = multiplier ? mult_desc->dtype : AP_F32;
    const int64_t* ms = multiplier ? mult_desc->strides : vol_desc->strides;
    // upright views go to the gather kernel (one pass when the workspace
    // holds every view, else view chunks that fit); the general kernel takes
    // the others
    int skip = 0;
    const uint32_t* left = nullptr;
    const uint64_t need = gather_workspace_bytes(sc.V, sc.nu, sc.nv, cfg->ssu, cfg->ssv);
    if (workspace && gather_supported(sc.nv, cfg->ssv) && need > 0) {
        uint32_t vc = sc.V;
        if (workspace_bytes < need) {
            vc = (uint32_t)((double)workspace_bytes / ((double)need / sc.V));
            while (vc > 0 && gather_workspace_bytes(vc, sc.nu, sc.nv, cfg->ssu, cfg->ssv) > workspace_bytes) vc--;
        }
        if (vc > 0) {
            const double dscale = (double)alpha * cfg->multiply_before_adding / ((double)cfg->ssu * cfg->ssv) *
                                  chunk->voxwid;
            const int64_t esz = dtype_size(sino_desc->dtype);
            for (uint32_t p0 = 0; p0 < sc.V; p0 += vc) {
                const uint32_t n = min(vc, sc.V - p0);
                st = launch_gather_adjoint(rc, n, sc.nu, sc.nv, sc.s0, sc.s1, sc.s2, cfg->ssu, cfg->ssv, dscale,
                                           specs + p0, vol_desc, vol, mdt, ms, multiplier, sino_desc->dtype,
                                           (const char*)sino + (int64_t)p0 * sc.s0 * esz, workspace,
                                           workspace_bytes, s, vc == sc.V ? &left : nullptr);
                if (st != AP_OK && st != AP_ERR_UNSUPPORTED) return st;
                if (st == AP_ERR_UNSUPPORTED) break;
            }
            skip = st == AP_OK ? 1 : 0;
            if (!skip) left = nullptr;
        }
    }
    return dispatch_dtype(vol_desc->dtype, [&](auto tv) -> int32_t {
        using Tv = decltype(tv);
        return dispatch_dtype(sino_desc->dtype, [&](auto tp) -> int32_t {
            using Tp = decltype(tp);
            return dispatch_dtype(mdt, [&](auto tm) -> int32_t {
                using Tm = decltype(tm);
                if constexpr (std::is_same<Tv, double>::value || std::is_same<Tp, double>::value ||
                              std::is_same<Tp, bf16_t>::value || std::is_same<Tm, double>::value) {
                    return ap_fail(AP_ERR_UNSUPPORTED, "unsupported dtype combination");
                } else {
                    if (idx64) {
                        hipLaunchKernelGGL((k_backproject_matched<Tv, Tp, Tm, int64_t>), grid, dim3(kAdjThreads),
                                           0, s, rc, sc, scale, cfg->ssu, cfg->ssv, specs, (Tv*)vol,
                                           vol_desc->strides[0], vol_desc->strides[1], vol_desc->strides[2],
                                           (const Tm*)multiplier, ms[0], ms[1], ms[2], (const Tp*)sino, nbx,
                                           nby, nbz, per_xcd, skip, left);
                    } else {
                        hipLaunchKernelGGL((k_backproject_matched<Tv, Tp, Tm, int32_t>), grid, dim3(kAdjThreads),
                                           0, s, rc, sc, scale, cfg->ssu, cfg->ssv, specs, (Tv*)vol,
                                           (int32_t)vol_desc->strides[0], (int32_t)vol_desc->strides[1],
                                           (int32_t)vol_desc->strides[2], (const Tm*)multiplier,
                                           (int32_t)ms[0], (int32_t)ms[1], (int32_t)ms[2], (const Tp*)sino,
                                           nbx, nby, nbz, per_xcd, skip, left);
                    }
                    return ap_check_launch("backproject_matched");
                }
            });
        });
    });
}

extern "C" uint64_t ap_backproject_matched_workspace_size(const ap_config_cb_project* cfg, uint32_t num_projs,
                                                          const ap_array3* sino_desc) {
    if (!cfg || !sino_desc || cfg->ssu < 1 || cfg->ssv < 1) return 0;
    if (!gather_supported(sino_desc->shape[2], cfg->ssv)) return 0;
    return gather_workspace_bytes(num_projs, sino_desc->shape[1], sino_desc->shape[2], cfg->ssu, cfg->ssv);
}
