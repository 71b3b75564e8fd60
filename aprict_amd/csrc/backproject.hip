// backproject.hip -- voxel-driven cone-beam backprojector with aprict's
// semantics (ap_backproject): the exact intersection-length branch and the
// fast interpolating (FDK / SIRT) branch of CB_backproject.cu:313-733,
// including sub-voxel sampling, cylinder mask, detector-edge softening,
// Z-power weights, the weighting presets of the direct ray-casting branch
// and the summand multiplier (:952-1027); and the GBC backprojector for a
// cylindrical source locus with its weight volume (:736-1108).
//
// One thread per voxel; a wavefront covers 64 consecutive voxels along the
// volume's z axis.  Views are processed in batches of kViewBatch: the block
// derives each view's projective map once (FP64, rounded to FP32) into LDS.
#include <cmath>

#include "ap_common.h"

namespace {

constexpr int kBz = 64;
// voxel rows (along y) per workgroup: AP_TUNE_BACKPROJECT variants; every
// workgroup has at least kViewBatch threads to stage the view maps
constexpr int kBpRowsY[3] = {4, 2, 1};
constexpr int kViewBatch = 64;
static_assert(kBz * kBpRowsY[2] >= kViewBatch, "every variant stages a full view batch");
constexpr int kGbcRows = 2;   // voxel rows per workgroup of the GBC kernel

// Per-view constants (FP32) of the central projection x -> detector:
//   cu = (Au.x + au)/(W.x + w0),  cv = (Av.x + av)/(W.x + w0)  (pixels from
//   the detector centre, geometry.py:75-107 expanded into a projective map),
//   lambda = K/(W.x + w0) is the source->detector / source->point ratio,
//   Z = Zc.x + z0 is the distance along the source->detector-centre axis.
struct ViewMap {
    float Au[3], au, Av[3], av, W[3], w0, K;
    float Zc[3], z0, Lnorm;
    float S[3], D[3], U[3], Vv[3];
    float nu, nv, uvnorm;
    float kq;              // K^2 / (|u| |v|): the FDK weight is kq w^3 / (W.x + w0)^2
    float hu, hv;          // (nu - 1) / 2, (nv - 1) / 2: pixel index of the detector centre
    float inv_hnu, inv_hnv;   // 2 / nu, 2 / nv
    float inv_Lnorm;          // 1 / |D - S|
};

__device__ void make_view_map(const ap_proj_spec* specs, uint32_t p, ViewMap& m) {
    const ViewGeom g = load_view(specs, p);
    const d3 n = cross(g.U, g.V);
    const double K = dot(g.D - g.S, n);
    const double uu = dot(g.U, g.U), vv = dot(g.V, g.V);
    const double a3u = dot(g.U, g.S - g.D), a3v = dot(g.V, g.S - g.D);
    const double nS = dot(n, g.S);
    const d3 Au = (n * a3u + g.U * K) * (1.0 / uu);
    const d3 Av = (n * a3v + g.V * K) * (1.0 / vv);
    m.Au[0] = (float)Au.x; m.Au[1] = (float)Au.y; m.Au[2] = (float)Au.z;
    m.au = (float)((-a3u * nS - K * dot(g.U, g.S)) / uu);
    m.Av[0] = (float)Av.x; m.Av[1] = (float)Av.y; m.Av[2] = (float)Av.z;
    m.av = (float)((-a3v * nS - K * dot(g.V, g.S)) / vv);
    m.W[0] = (float)n.x; m.W[1] = (float)n.y; m.W[2] = (float)n.z;
    m.w0 = (float)(-nS);
    m.K = (float)K;
    const d3 L = g.D - g.S;
    const double ln = sqrt(dot(L, L));
    m.Zc[0] = (float)(L.x / ln); m.Zc[1] = (float)(L.y / ln); m.Zc[2] = (float)(L.z / ln);
    m.z0 = (float)(-dot(g.S, L) / ln);
    m.Lnorm = (float)ln;
    m.inv_Lnorm = (float)(1.0 / ln);
    m.S[0] = (float)g.S.x; m.S[1] = (float)g.S.y; m.S[2] = (float)g.S.z;
    m.D[0] = (float)g.D.x; m.D[1] = (float)g.D.y; m.D[2] = (float)g.D.z;
    m.U[0] = (float)g.U.x; m.U[1] = (float)g.U.y; m.U[2] = (float)g.U.z;
    m.Vv[0] = (float)g.V.x; m.Vv[1] = (float)g.V.y; m.Vv[2] = (float)g.V.z;
    m.nu = (float)g.nu;
    m.nv = (float)g.nv;
    m.uvnorm = (float)(sqrt(uu) * sqrt(vv));
    m.kq = (float)(K * K / (sqrt(uu) * sqrt(vv)));
    m.hu = (float)(0.5 * ((double)g.nu - 1.0));
    m.hv = (float)(0.5 * ((double)g.nv - 1.0));
    m.inv_hnu = (float)(2.0 / (double)g.nu);
    m.inv_hnv = (float)(2.0 / (double)g.nv);
}

// cos^2-type softening window s(a, b; t) of the GBC paper's appendix
// (CB_backproject.cu:274-311), 1 for |t| < a - b, 0 for |t| > a and
// 3x^2 - 2x^3 between with x = (sin|t| - sin a) / (sin(a - b) - sin a), held as
// constants of sin|t|.  Both GBC angles (the elevation and the horizontal
// angle) lie in [0, pi/2], where sine increases, so |t| > a <=> sin|t| > hi
// and |t| < a - b <=> sin|t| < lo, with hi = sin a (2 when a >= pi/2: never
// exceeded, -1 when a < 0: always) and lo = sin(a - b) (-1 when a - b <= 0:
// never undercut, 2 when a - b > pi/2: always); g = 1 / (sin(a - b) - sin a)
// and gs = g sin a.
struct SoftWin {
    float hi, lo, g, gs;
};

// GBC weighting constants: 1 / R (R the radius of the source cylinder about
// the z axis) and the windows of the elevation (vartheta_v) and of the
// horizontal angle (vartheta_h)
struct GbcSoft {
    float inv_R;
    SoftWin v, h;
};

inline SoftWin make_soft_win(float a32, float b32) {
    const double a = a32, d = (double)a32 - (double)b32, hp = 1.5707963267948966;
    SoftWin s;
    s.hi = a < 0.0 ? -1.f : (a >= hp ? 2.f : (float)std::sin(a));
    s.lo = d <= 0.0 ? -1.f : (d > hp ? 2.f : (float)std::sin(d));
    const double g = 1.0 / (std::sin(d) - std::sin(a));
    s.g = (float)g;
    s.gs = (float)(g * std::sin(a));
    return s;
}

inline GbcSoft make_gbc_soft(const ap_config_cb_backproject* cfg, double R) {
    return GbcSoft{(float)(1.0 / R), make_soft_win(cfg->soft_vartheta_v, cfg->soft_vartheta_v_softspan),
                   make_soft_win(cfg->soft_vartheta_h, cfg->soft_vartheta_h_softspan)};
}

struct BpCfg {
    bool exact, all1s, mask, direct, bilinear, use_mult;
    int preset;           // backproj_weighting_preset
    GbcSoft gbc;          // GBC weighting (preset 7)
    float mask_r2, sep, zval;
    float inv_zval;       // 1 / zval
    float inv_sep;        // 1 / sep (0 when sep == 0)
    float w3;             // voxwid^3
    bool off32;           // detector offsets of one view fit 32 bits
    bool pair;            // v contiguous (s2 == 1), nv >= 2, 32-bit offsets: rows k, k+1 in one load
    bool grouped;         // bp_col_views over 4 views applies: offsets of 4 views fit 32 bits (and pairs if bilinear)
    int zl, zv;
    uint32_t ssx, ssy, ssz;
    float w;              // voxwid
    float c000[3];        // centre of voxel (0,0,0) of the whole volume
    uint32_t off[3], shp[3];
    uint32_t nu, nv;      // sinogram shape
    int64_t s0, s1, s2;   // sinogram strides
};

// b^e for an integer exponent: one hardware reciprocal (about 1 ulp; an IEEE
// division costs ten instructions per voxel-view) for negative powers, then
// products
__device__ __forceinline__ float powi(float b, int e) {
    const float f = e < 0 ? __builtin_amdgcn_rcpf(b) : b;
    float r = 1.f;
    for (int i = 0, n = e < 0 ? -e : e; i < n; i++) r *= f;
    return r;
}

template <typename Tp>
__device__ __forceinline__ float det_read(const Tp* __restrict__ img, const BpCfg& c, int j, int k) {
    if (j < 0 || k < 0 || j >= (int)c.nu || k >= (int)c.nv) return 0.f;
    return to_f32(img[(int64_t)j * c.s1 + (int64_t)k * c.s2]);
}

// length of the ray from S through P (unbounded beyond P) inside the cube of
// side w centred on C, CB_backproject.cu:96-131
__device__ __forceinline__ float ray_cube_len(const float S[3], const float P[3],
                                              const float C[3], float w) {
    float t0 = 0.f, t1 = AP_BIG, d2 = 0.f;
    for (int a = 0; a < 3; a++) {
        const float d = P[a] - S[a];
        d2 += d * d;
        const float lo = C[a] - 0.5f * w - S[a], hi = C[a] + 0.5f * w - S[a];
        if (d == 0.f) {
            if (lo > 0.f || hi < 0.f) return 0.f;
            continue;
        }
        const float inv = 1.f / d;
        const float ta = lo * inv, tb = hi * inv;
        t0 = fmaxf(t0, fminf(ta, tb));
        t1 = fminf(t1, fmaxf(ta, tb));
    }
    return t1 > t0 ? (t1 - t0) * sqrtf(d2) : 0.f;
}

// Detector-edge softening (CB_backproject.cu:694-713): a cosine ramp over the
// outer `sep` fraction of each half of the detector.  Outside that band both
// factors are exactly 1 (cos 0 = 1), so the cosines are evaluated only for
// voxels that project into it.
__device__ __forceinline__ float soften(const BpCfg& c, const ViewMap& m, float cu, float cv, float val) {
    if (!(c.sep > 0.f) || val == 0.f) return val;
    float ud = fabsf(cu) * m.inv_hnu, vd = fabsf(cv) * m.inv_hnv;
    ud = fmaxf(0.f, fminf(1.f, (1.f - ud) * c.inv_sep));
    vd = fmaxf(0.f, fminf(1.f, (1.f - vd) * c.inv_sep));
    if (ud < 1.f || vd < 1.f) {
        const float pi = 3.14159265358979f;
        ud = fmaxf(0.f, fminf(1.f, 0.5f * (1.f + cosf(pi * (1.f - ud)))));
        vd = fmaxf(0.f, fminf(1.f, 0.5f * (1.f + cosf(pi * (1.f - vd)))));
        val *= ud * vd;
    }
    return val;
}

// Weight of the direct ray-casting branch for backproj_weighting_preset 1-4
// and 6 (CB_backproject.cu:528-600), applied to `mult`.  R is Z_on_val_factor_val, the
// source locus is a cylinder about the z axis and zz the distance from the
// source to the voxel along the source->detector-centre axis.  Other presets
// add no weight there (preset 7 has its own kernel).
__device__ __forceinline__ float apply_preset(const BpCfg& c, const ViewMap& m, const float x[3], float zz,
                                              float mult) {
    const float R = c.zval;
    const float r0 = x[0] - m.S[0], r1 = x[1] - m.S[1];
    const float ell = sqrtf(r0 * r0 + r1 * r1) / R;   // horizontal source->voxel distance / R
    const float zr = zz / R;
    const float cr = fmaxf(0.f, fminf(1.f, fabsf(zr / ell)));
    const float dr = cr - ell;
    // voxel position relative to the source direction, in units of R^2
    const float vy = (m.S[0] * x[1] - m.S[1] * x[0]) / R / R;
    const float vx = (m.S[0] * x[0] + m.S[1] * x[1]) / R / R;
    switch (c.preset) {
    case 1:
        return mult / (cr * cr * (cr * cr + dr * dr));
    case 2: {
        const float q = dr / cr;
        const float f = (1.f - q * q) / (1.f + q * q);
        return mult * (-2.f * (f * f) * (-vy) * (1.f + vx / zr) / (ell * ell));
    }
    case 3:
        return mult / (zr * cr * cr);
    case 4:
        return mult * ((vy / zr) / zr);
    case 6: {
        const float q = dr / cr;
        const float f = (1.f - q * q) / (1.f + q * q);
        return mult * (-2.f * (f * f) / (1.f - q) * (-vy) * (1.f + vx / zr) / zr);
    }
    default:
        return mult;
    }
}

template <typename Tp>
__device__ __forceinline__ float sample_det(const BpCfg& c, const ViewMap& m, const Tp* __restrict__ img,
                                            float cu, float cv);

// The interpolating branch (CB_backproject.cu:478-691): the voxel centre's
// projection, weighted by the FDK magnification factor (or 1 with direct
// ray-casting) and Z powers, read bilinearly or at the nearest pixel.  The
// reciprocal of the projective denominator is the hardware one (within an
// ulp of a division).  Taps outside the detector read 0.
template <typename Tp>
__device__ __forceinline__ float bp_fast_view(const BpCfg& c, const ViewMap& m, const Tp* __restrict__ img,
                                              const float x[3]) {
    const float wx = m.W[0] * x[0] + m.W[1] * x[1] + m.W[2] * x[2] + m.w0;
    const float iw = __builtin_amdgcn_rcpf(wx);
    const float cu = (m.Au[0] * x[0] + m.Au[1] * x[1] + m.Au[2] * x[2] + m.au) * iw;
    const float cv = (m.Av[0] * x[0] + m.Av[1] * x[1] + m.Av[2] * x[2] + m.av) * iw;
    float mult = c.direct ? 1.f : m.kq * c.w3 * (iw * iw);
    const bool pre = c.direct && c.preset != 0;
    if (c.zl != 0 || c.zv != 0 || pre) {
        const float zz = m.Zc[0] * x[0] + m.Zc[1] * x[1] + m.Zc[2] * x[2] + m.z0;
        if (c.zl != 0) mult *= powi(zz * m.inv_Lnorm, c.zl);
        if (c.zv != 0) mult *= powi(zz * c.inv_zval, c.zv);
        if (pre) mult = apply_preset(c, m, x, zz, mult);
    }
    if (mult == 0.f) return 0.f;
    return soften(c, m, cu, cv, sample_det(c, m, img, cu, cv) * mult);
}

// two detector values adjacent along v, read in one load (element alignment)
template <typename Tp>
struct alignas(sizeof(Tp)) TapPair {
    Tp lo, hi;
};

// Detector value at (cu, cv) pixels from the centre: bilinear
// (CB_backproject.cu:221-254) or nearest (:256-271), zero off the detector.
template <typename Tp>
__device__ __forceinline__ float sample_det(const BpCfg& c, const ViewMap& m, const Tp* __restrict__ img,
                                            float cu, float cv) {
    auto rd = [&](int j, int k) -> float {
        if (c.off32) return to_f32(img[j * (int32_t)c.s1 + k * (int32_t)c.s2]);
        return to_f32(img[(int64_t)j * c.s1 + (int64_t)k * c.s2]);
    };
    const float fj = cu + m.hu, fk = cv + m.hv;
    float y = 0.f;
    if (c.bilinear) {
        if (fj > -1.f && fk > -1.f && fj < (float)c.nu && fk < (float)c.nv) {
            const float fl_j = floorf(fj), fl_k = floorf(fk);
            const int j0 = (int)fl_j, k0 = (int)fl_k;
            const float a = fj - fl_j, b = fk - fl_k;
            if (c.pair) {
                // rows k0, k0 + 1 of columns ja, jb from the pair starting at
                // kc = clamp(k0, 0, nv - 2): kc = k0 inside the detector, and at
                // its edges (k0 = -1 or nv - 1) the missing row reads 0
                const int ja = max(j0, 0), jb = min(j0 + 1, (int)c.nu - 1);
                const int kc = min(max(k0, 0), (int)c.nv - 2);
                const TapPair<Tp> pa = *reinterpret_cast<const TapPair<Tp>*>(img + (ja * (int32_t)c.s1 + kc));
                const TapPair<Tp> pb = *reinterpret_cast<const TapPair<Tp>*>(img + (jb * (int32_t)c.s1 + kc));
                const bool klo = k0 < 0, khi = k0 + 1 >= (int)c.nv;
                if (__ballot(klo || khi || j0 < 0 || j0 + 1 >= (int)c.nu) == 0) {
                    // no tap of the wave's samples leaves the detector: the same
                    // blend without the edge selects
                    const float t1 = to_f32(pa.lo) * (1.f - a) + to_f32(pb.lo) * a;
                    const float t2 = to_f32(pa.hi) * (1.f - a) + to_f32(pb.hi) * a;
                    return t1 * (1.f - b) + t2 * b;
                }
                float y00 = klo ? 0.f : (khi ? to_f32(pa.hi) : to_f32(pa.lo));
                float y01 = khi ? 0.f : (klo ? to_f32(pa.lo) : to_f32(pa.hi));
                float y10 = klo ? 0.f : (khi ? to_f32(pb.hi) : to_f32(pb.lo));
                float y11 = khi ? 0.f : (klo ? to_f32(pb.lo) : to_f32(pb.hi));
                if (j0 < 0) { y00 = 0.f; y01 = 0.f; }
                if (j0 + 1 >= (int)c.nu) { y10 = 0.f; y11 = 0.f; }
                const float t1 = y00 * (1.f - a) + y10 * a;
                const float t2 = y01 * (1.f - a) + y11 * a;
                return t1 * (1.f - b) + t2 * b;
            }
            // the four taps at clamped (in-range) indices, zeroed off the detector
            const int ja = max(j0, 0), jb = min(j0 + 1, (int)c.nu - 1);
            const int ka = max(k0, 0), kb = min(k0 + 1, (int)c.nv - 1);
            float y00 = rd(ja, ka), y10 = rd(jb, ka), y01 = rd(ja, kb), y11 = rd(jb, kb);
            if (j0 < 0) { y00 = 0.f; y01 = 0.f; }
            if (j0 + 1 >= (int)c.nu) { y10 = 0.f; y11 = 0.f; }
            if (k0 < 0) { y00 = 0.f; y10 = 0.f; }
            if (k0 + 1 >= (int)c.nv) { y01 = 0.f; y11 = 0.f; }
            const float t1 = y00 * (1.f - a) + y10 * a;
            const float t2 = y01 * (1.f - a) + y11 * a;
            y = t1 * (1.f - b) + t2 * b;
        }
    } else {
        const int j = (int)floorf(fj + 0.5f), k = (int)floorf(fk + 0.5f);
        if (j >= 0 && k >= 0 && j < (int)c.nu && k < (int)c.nv) y = rd(j, k);
    }
    return y;
}

// A view is z-coherent when its projective map does not depend on z except
// through cv: W.z = 0 and Au.z = 0 (detector normal and u axis horizontal),
// and, where the Z weights or a weighting preset apply, Zc.z = 0 (horizontal
// source->detector axis).  Along a voxel column the denominator, cu, the
// u taps and the weight are then constant and cv is linear in z.
__device__ __forceinline__ bool view_z_coherent(const BpCfg& c, const ViewMap& m) {
    const bool zdep = c.zl != 0 || c.zv != 0 || (c.direct && c.preset != 0);
    return m.W[2] == 0.f && m.Au[2] == 0.f && (!zdep || m.Zc[2] == 0.f);
}

// The interpolating branch for kV consecutive views at kZ voxels of one
// column (x, y) at heights z[m], all detector reads issued before the
// blends (kV kZ, bilinear 2 kV kZ, gathers in flight per lane).  With kCoh
// (every view z-coherent) a view's u part -- denominator, cu, the weight, the
// u taps -- is formed once and cv = cv(0) + z dcv per voxel; otherwise each
// voxel forms the whole map as bp_fast_view does.  The blends and weights are
// those of bp_fast_view / sample_det, and each voxel adds its views in order,
// so a voxel's sum does not depend on kV or kZ.  Reads use clamped in-range
// addresses and are zeroed where sample_det reads nothing.  Needs 32-bit
// offsets over the kV views (view v at img + v s0), and for bilinear taps the
// paired layout (c.pair).
template <int kV, int kZ, bool kCoh, typename Tp>
__device__ __forceinline__ void bp_col_views(const BpCfg& c, const ViewMap* Ms, const Tp* __restrict__ img,
                                             int64_t s0, float x0, float x1, const float (&z)[kZ],
                                             uint32_t act, float (&acc)[kZ]) {
    float cu[kV][kZ], cv[kV][kZ], mult[kV][kZ], y[kV][kZ];
#pragma unroll
    for (int v = 0; v < kV; v++) {
        const ViewMap& M = Ms[v];
        const bool pre = c.direct && c.preset != 0;
        if constexpr (kCoh) {
            const float wx = M.W[0] * x0 + M.W[1] * x1 + M.w0;
            const float iw = __builtin_amdgcn_rcpf(wx);
            const float cuv = (M.Au[0] * x0 + M.Au[1] * x1 + M.au) * iw;
            const float cv0 = (M.Av[0] * x0 + M.Av[1] * x1 + M.av) * iw;
            const float dcv = M.Av[2] * iw;
            float mu = c.direct ? 1.f : M.kq * c.w3 * (iw * iw);
            if (c.zl != 0 || c.zv != 0 || pre) {
                const float zz = M.Zc[0] * x0 + M.Zc[1] * x1 + M.z0;
                if (c.zl != 0) mu *= powi(zz * M.inv_Lnorm, c.zl);
                if (c.zv != 0) mu *= powi(zz * c.inv_zval, c.zv);
                if (pre) {
                    const float xp[3] = {x0, x1, 0.f};   // the presets use x and y only
                    mu = apply_preset(c, M, xp, zz, mu);
                }
            }
#pragma unroll
            for (int m = 0; m < kZ; m++) {
                cu[v][m] = cuv;
                cv[v][m] = __builtin_fmaf(z[m], dcv, cv0);
                mult[v][m] = mu;
            }
        } else {
#pragma unroll
            for (int m = 0; m < kZ; m++) {
                const float x[3] = {x0, x1, z[m]};
                const float wx = M.W[0] * x[0] + M.W[1] * x[1] + M.W[2] * x[2] + M.w0;
                const float iw = __builtin_amdgcn_rcpf(wx);
                cu[v][m] = (M.Au[0] * x[0] + M.Au[1] * x[1] + M.Au[2] * x[2] + M.au) * iw;
                cv[v][m] = (M.Av[0] * x[0] + M.Av[1] * x[1] + M.Av[2] * x[2] + M.av) * iw;
                float mu = c.direct ? 1.f : M.kq * c.w3 * (iw * iw);
                if (c.zl != 0 || c.zv != 0 || pre) {
                    const float zz = M.Zc[0] * x[0] + M.Zc[1] * x[1] + M.Zc[2] * x[2] + M.z0;
                    if (c.zl != 0) mu *= powi(zz * M.inv_Lnorm, c.zl);
                    if (c.zv != 0) mu *= powi(zz * c.inv_zval, c.zv);
                    if (pre) mu = apply_preset(c, M, x, zz, mu);
                }
                mult[v][m] = mu;
            }
        }
    }
    if (c.bilinear) {
        TapPair<Tp> pa[kV][kZ], pb[kV][kZ];
        float ta[kV][kZ], tb[kV][kZ];
        // tap indices and the in-range flag for the blend with edges.  Per-voxel
        // maps keep them as integers (opaque copies): as flags across the branch
        // they would be lane masks, held in scalar registers; a z-coherent view
        // shares its u taps and needs fewer
        int j0v[kV][kZ], k0v[kV][kZ], inv[kV][kZ];
        bool edge = false;
#pragma unroll
        for (int v = 0; v < kV; v++) {
            const ViewMap& M = Ms[v];
            const int vo = v * (int32_t)s0;
            // a z-coherent view's u taps, once: where the voxel's sample is in
            // range they equal the per-voxel ones, and elsewhere it reads 0
            const float fju = cu[v][0] + M.hu;
            const bool uin = fju > -1.f && fju < (float)c.nu;
            const float fjcu = uin ? fju : 0.f;
            const float flu = floorf(fjcu);
            const int j0u = (int)flu;
            const float tau = fjcu - flu;
#pragma unroll
            for (int m = 0; m < kZ; m++) {
                const float fk = cv[v][m] + M.hv;
                bool inr;
                int j0;
                if constexpr (kCoh) {
                    inr = uin && fk > -1.f && fk < (float)c.nv;
                    j0 = j0u;
                    ta[v][m] = tau;
                } else {
                    const float fj = cu[v][m] + M.hu;
                    inr = fj > -1.f && fk > -1.f && fj < (float)c.nu && fk < (float)c.nv;
                    const float fjc = inr ? fj : 0.f;   // finite before the conversion
                    const float fl_j = floorf(fjc);
                    j0 = (int)fl_j;
                    ta[v][m] = fjc - fl_j;
                }
                const float fkc = inr ? fk : 0.f;
                const float fl_k = floorf(fkc);
                const int k0 = (int)fl_k;
                tb[v][m] = fkc - fl_k;
                const int ja = max(j0, 0), jb = min(j0 + 1, (int)c.nu - 1);
                const int kc = min(max(k0, 0), (int)c.nv - 2);
                pa[v][m] = *reinterpret_cast<const TapPair<Tp>*>(img + (vo + ja * (int32_t)c.s1 + kc));
                pb[v][m] = *reinterpret_cast<const TapPair<Tp>*>(img + (vo + jb * (int32_t)c.s1 + kc));
                edge = edge || (((act >> m) & 1u) &&
                                (!inr || k0 < 0 || k0 + 1 >= (int)c.nv || j0 < 0 || j0 + 1 >= (int)c.nu));
                j0v[v][m] = j0;
                k0v[v][m] = k0;
                inv[v][m] = inr ? 1 : 0;
                if constexpr (!kCoh) asm volatile("" : "+v"(j0v[v][m]), "+v"(k0v[v][m]), "+v"(inv[v][m]));
            }
        }
        if (__ballot(edge) == 0) {
            // no tap of the wave's active voxels falls off the detector: the blend of
            // the two loaded pairs without the edge selects (the same values)
#pragma unroll
            for (int v = 0; v < kV; v++)
#pragma unroll
                for (int m = 0; m < kZ; m++) {
                    const float a = ta[v][m], b = tb[v][m];
                    const float t1 = to_f32(pa[v][m].lo) * (1.f - a) + to_f32(pb[v][m].lo) * a;
                    const float t2 = to_f32(pa[v][m].hi) * (1.f - a) + to_f32(pb[v][m].hi) * a;
                    y[v][m] = t1 * (1.f - b) + t2 * b;
                }
        } else {
#pragma unroll
            for (int v = 0; v < kV; v++)
#pragma unroll
                for (int m = 0; m < kZ; m++) {
                    y[v][m] = 0.f;
                    if (inv[v][m] != 0) {
                        const int j0 = j0v[v][m], k0 = k0v[v][m];
                        const bool lo = k0 < 0, hi = k0 + 1 >= (int)c.nv;
                        float y00 = lo ? 0.f : (hi ? to_f32(pa[v][m].hi) : to_f32(pa[v][m].lo));
                        float y01 = hi ? 0.f : (lo ? to_f32(pa[v][m].lo) : to_f32(pa[v][m].hi));
                        float y10 = lo ? 0.f : (hi ? to_f32(pb[v][m].hi) : to_f32(pb[v][m].lo));
                        float y11 = hi ? 0.f : (lo ? to_f32(pb[v][m].lo) : to_f32(pb[v][m].hi));
                        if (j0 < 0) { y00 = 0.f; y01 = 0.f; }
                        if (j0 + 1 >= (int)c.nu) { y10 = 0.f; y11 = 0.f; }
                        const float a = ta[v][m], b = tb[v][m];
                        const float t1 = y00 * (1.f - a) + y10 * a;
                        const float t2 = y01 * (1.f - a) + y11 * a;
                        y[v][m] = t1 * (1.f - b) + t2 * b;
                    }
                }
        }
    } else {
        Tp yv[kV][kZ];
        bool inr[kV][kZ];
#pragma unroll
        for (int v = 0; v < kV; v++) {
            const ViewMap& M = Ms[v];
            const int vo = v * (int32_t)s0;
#pragma unroll
            for (int m = 0; m < kZ; m++) {
                // nearest: the index clamped to [-1, n] keeps the conversion defined
                const float fj = cu[v][m] + M.hu, fk = cv[v][m] + M.hv;
                const int j = (int)floorf(fminf(fmaxf(fj + 0.5f, -1.f), (float)c.nu));
                const int k = (int)floorf(fminf(fmaxf(fk + 0.5f, -1.f), (float)c.nv));
                inr[v][m] = j >= 0 && k >= 0 && j < (int)c.nu && k < (int)c.nv;
                const int jc = min(max(j, 0), (int)c.nu - 1), kk = min(max(k, 0), (int)c.nv - 1);
                yv[v][m] = img[vo + jc * (int32_t)c.s1 + kk * (int32_t)c.s2];
            }
        }
#pragma unroll
        for (int v = 0; v < kV; v++)
#pragma unroll
            for (int m = 0; m < kZ; m++) y[v][m] = inr[v][m] ? to_f32(yv[v][m]) : 0.f;
    }
#pragma unroll
    for (int v = 0; v < kV; v++)
#pragma unroll
        for (int m = 0; m < kZ; m++)
            acc[m] += mult[v][m] == 0.f ? 0.f : soften(c, Ms[v], cu[v][m], cv[v][m], y[v][m] * mult[v][m]);
}

template <typename Tp>
__device__ float bp_one_view(const BpCfg& c, const ViewMap& m, const Tp* __restrict__ img,
                             const float x[3]) {
    if (!c.exact) return bp_fast_view(c, m, img, x);
    const float wx = m.W[0] * x[0] + m.W[1] * x[1] + m.W[2] * x[2] + m.w0;
    const float iw = 1.f / wx;
    const float cu = (m.Au[0] * x[0] + m.Au[1] * x[1] + m.Au[2] * x[2] + m.au) * iw;
    const float cv = (m.Av[0] * x[0] + m.Av[1] * x[1] + m.Av[2] * x[2] + m.av) * iw;
    float val = 0.f;
    if (c.exact) {
        float mnu = cu, mxu = cu, mnv = cv, mxv = cv;
        for (int q = 0; q < 8; q++) {
            const float y0 = x[0] + ((q & 4) ? 0.5f : -0.5f) * c.w;
            const float y1 = x[1] + ((q & 2) ? 0.5f : -0.5f) * c.w;
            const float y2 = x[2] + ((q & 1) ? 0.5f : -0.5f) * c.w;
            const float iw2 = 1.f / (m.W[0] * y0 + m.W[1] * y1 + m.W[2] * y2 + m.w0);
            const float a = (m.Au[0] * y0 + m.Au[1] * y1 + m.Au[2] * y2 + m.au) * iw2;
            const float b = (m.Av[0] * y0 + m.Av[1] * y1 + m.Av[2] * y2 + m.av) * iw2;
            mnu = fminf(mnu, a); mxu = fmaxf(mxu, a);
            mnv = fminf(mnv, b); mxv = fmaxf(mxv, b);
        }
        const int j0 = max(0, (int)floorf(mnu + 0.5f * m.nu));
        const int k0 = max(0, (int)floorf(mnv + 0.5f * m.nv));
        const int j1 = min((int)c.nu - 1, (int)ceilf(mxu + 0.5f * m.nu));
        const int k1 = min((int)c.nv - 1, (int)ceilf(mxv + 0.5f * m.nv));
        for (int j = j0; j <= j1; j++) {
            const float a = (float)j - 0.5f * (m.nu - 1.f);
            for (int k = k0; k <= k1; k++) {
                const float b = (float)k - 0.5f * (m.nv - 1.f);
                const float P[3] = {m.D[0] + a * m.U[0] + b * m.Vv[0],
                                    m.D[1] + a * m.U[1] + b * m.Vv[1],
                                    m.D[2] + a * m.U[2] + b * m.Vv[2]};
                const float len = ray_cube_len(m.S, P, x, c.w);
                if (len != 0.f) {
                    const float y = c.all1s ? 1.f : to_f32(img[(int64_t)j * c.s1 + (int64_t)k * c.s2]);
                    val += y * len;
                }
            }
        }
        if (c.zl != 0) {
            const float zz = m.Zc[0] * x[0] + m.Zc[1] * x[1] + m.Zc[2] * x[2] + m.z0;
            val *= powi(zz * m.inv_Lnorm, c.zl);
        }
        if (c.zv != 0) {
            const float zz = m.Zc[0] * x[0] + m.Zc[1] * x[1] + m.Zc[2] * x[2] + m.z0;
            val *= powi(zz * c.inv_zval, c.zv);
        }
    }
    return soften(c, m, cu, cv, val);
}

// One thread per kZ voxels of a column, kBz apart along z (a wave holds kBz
// consecutive voxels at each of its kZ heights).  The interpolating branch
// without sub-voxel sampling runs bp_col_views, with the shared u part on a
// batch of views that are all z-coherent; the host takes kZ > 1 only there.
// A voxel's sum is the same for every kZ and launch shape.
template <int kBy, int kZ, typename Tv, typename Tp, typename Tm>
__global__ __launch_bounds__(kBz* kBy) void k_backproject_voxel(
    BpCfg c, const ap_proj_spec* __restrict__ specs, uint32_t num_projs,
    Tv* __restrict__ vol, int64_t vs0, int64_t vs1, int64_t vs2,
    const Tm* __restrict__ mult, int64_t ms0, int64_t ms1, int64_t ms2,
    const Tp* __restrict__ sino) {
    __shared__ ViewMap maps[kViewBatch];
    const uint32_t ck0 = blockIdx.x * (kBz * kZ) + threadIdx.x;
    const uint32_t cj = blockIdx.y * kBy + threadIdx.y;
    const uint32_t ci = blockIdx.z;
    const uint32_t gi = ci + c.off[0], gj = cj + c.off[1];
    const float x0 = c.c000[0] + (float)gi * c.w, x1 = c.c000[1] + (float)gj * c.w;
    bool col_in = cj < c.shp[1] && ci < c.shp[0];
    if (c.mask && x0 * x0 + x1 * x1 > c.mask_r2) col_in = false;
    // the thread's voxels in the volume: bit m for voxel m (an integer, not
    // flags: flags live across the view loop would occupy scalar registers)
    float z[kZ], acc[kZ];
    uint32_t act = 0;
#pragma unroll
    for (int m = 0; m < kZ; m++) {
        const uint32_t ck = ck0 + (uint32_t)(kBz * m);
        z[m] = c.c000[2] + (float)(ck + c.off[2]) * c.w;
        if (col_in && ck < c.shp[2]) act |= 1u << m;
        acc[m] = 0.f;
    }
    asm volatile("" : "+v"(act));
    const bool any = act != 0;
    const uint32_t nss = c.ssx * c.ssy * c.ssz;
    const bool zcol = !c.exact && nss == 1 && (c.bilinear ? c.pair : c.off32);
    const int tid = threadIdx.y * kBz + threadIdx.x;
    for (uint32_t p0 = 0; p0 < num_projs; p0 += kViewBatch) {
        const uint32_t nb = min((uint32_t)kViewBatch, num_projs - p0);
        __syncthreads();
        bool coherent = true;
        if (tid < (int)nb) {
            make_view_map(specs, p0 + tid, maps[tid]);
            coherent = view_z_coherent(c, maps[tid]);
        }
        // every view of the batch z-coherent (the barrier publishes the maps)
        const bool batch_z = __syncthreads_and(coherent) != 0;
        if (!any || nss != 1) continue;
        if (zcol) {
            const Tp* img = sino + (int64_t)p0 * c.s0;
            uint32_t q = 0;
            if constexpr (kZ == 1) {
                if (c.grouped) {
                    for (; q + 4 <= nb; q += 4) {
                        if (batch_z)
                            bp_col_views<4, 1, true>(c, maps + q, img + (int64_t)q * c.s0, c.s0, x0, x1, z, act, acc);
                        else
                            bp_col_views<4, 1, false>(c, maps + q, img + (int64_t)q * c.s0, c.s0, x0, x1, z, act, acc);
                    }
                }
            }
            for (; q < nb; q++) {
                if (batch_z)
                    bp_col_views<1, kZ, true>(c, maps + q, img + (int64_t)q * c.s0, c.s0, x0, x1, z, act, acc);
                else
                    bp_col_views<1, kZ, false>(c, maps + q, img + (int64_t)q * c.s0, c.s0, x0, x1, z, act, acc);
            }
            continue;
        }
        if constexpr (kZ == 1) {
            // the exact branch, and the interpolating branch with unpaired taps or
            // 64-bit detector offsets, voxel by voxel
            const float xc[3] = {x0, x1, z[0]};
            if (c.exact) {
                for (uint32_t q = 0; q < nb; q++)
                    acc[0] += bp_one_view(c, maps[q], sino + (int64_t)(p0 + q) * c.s0, xc);
            } else {
                for (uint32_t q = 0; q < nb; q++)
                    acc[0] += bp_fast_view(c, maps[q], sino + (int64_t)(p0 + q) * c.s0, xc);
            }
        }
    }
    if (nss != 1) {
        // general sub-voxel sampling (kZ = 1 only): views re-derived per sub-voxel
        // (CB_backproject.cu:377-383 loop order: x, y, z outermost to innermost)
#pragma unroll
        for (int m = 0; m < kZ; m++) {
            if (!((act >> m) & 1u)) continue;
            float ax = 0.f;
            for (uint32_t si = 0; si < c.ssx; si++) {
                float ay = 0.f;
                for (uint32_t sj = 0; sj < c.ssy; sj++) {
                    float az = 0.f;
                    for (uint32_t sk = 0; sk < c.ssz; sk++) {
                        const float xs[3] = {x0 + (-0.5f + (0.5f + si) / c.ssx) * c.w,
                                             x1 + (-0.5f + (0.5f + sj) / c.ssy) * c.w,
                                             z[m] + (-0.5f + (0.5f + sk) / c.ssz) * c.w};
                        float a = 0.f;
                        for (uint32_t p = 0; p < num_projs; p++) {
                            ViewMap vm;
                            make_view_map(specs, p, vm);
                            a += bp_one_view(c, vm, sino + (int64_t)p * c.s0, xs);
                        }
                        az += a;
                    }
                    ay += az / c.ssz;
                }
                ax += ay / c.ssy;
            }
            acc[m] = ax / c.ssx;
        }
    }
#pragma unroll
    for (int m = 0; m < kZ; m++) {
        if (!((act >> m) & 1u) || acc[m] == 0.f) continue;
        const int64_t gk = (int64_t)(ck0 + (uint32_t)(kBz * m) + c.off[2]);
        float mval = 1.f;
        if (c.use_mult) mval = to_f32(mult[(int64_t)gi * ms0 + (int64_t)gj * ms1 + gk * ms2]);
        const int64_t o = (int64_t)gi * vs0 + (int64_t)gj * vs1 + gk * vs2;
        vol[o] = from_f32<Tv>(to_f32(vol[o]) + acc[m] * mval);
    }
}

// the window s at sn = sin|t|
__device__ __forceinline__ float soft_win(const SoftWin& s, float sn) {
    if (sn > s.hi) return 0.f;
    if (sn < s.lo) return 1.f;
    const float x = s.g * sn - s.gs;
    return fmaxf(0.f, fminf(1.f, 3.f * x * x - 2.f * x * x * x));
}

// GBC weight of one view for a source S on a cylinder of radius R about the z
// axis (CB_backproject.cu:736-790, eq. (23) of the GBC paper): with ell the
// horizontal source->voxel distance over R, c the |cosine| of the horizontal
// angle between that direction and the direction to the axis, and theta the
// elevation of the voxel seen from the source,
//   w = cos^3(theta) c / (c^2 + (c - ell)^2) * s(v; theta) * s(h; acos c).
// Trigonometry by square roots: with r = x - S, cos theta = |r_xy| / |r|,
// sin|theta| = |r_z| / |r|, c = |r_xy . S_xy| / (|r_xy| |S_xy|) and
// sin(acos c) = sqrt(1 - c^2).  A point straight above or below the source
// (r_xy = 0) gets c = 1 and cos theta = 0, as the angle form gives.
__device__ __forceinline__ float gbc_weight(const GbcSoft& g, const float S[3], const float x[3]) {
    const float r0 = x[0] - S[0], r1 = x[1] - S[1], r2 = x[2] - S[2];
    const float rxy2 = r0 * r0 + r1 * r1;
    const float rxy = sqrtf(rxy2);
    const float ir = rsqrtf(rxy2 + r2 * r2);
    const float il = rsqrtf(S[0] * S[0] + S[1] * S[1]);
    const float cr = fminf(1.f, fabsf(r0 * S[0] + r1 * S[1]) * il * rsqrtf(rxy2));
    const float dr = cr - rxy * g.inv_R;
    const float ce = rxy * ir;
    // hardware reciprocal (about 1 ulp) instead of an IEEE division
    const float w = ce * ce * ce * cr * __builtin_amdgcn_rcpf(fmaxf(1e-9f, cr * cr + dr * dr));
    return w * (soft_win(g.v, fabsf(r2) * ir) * soft_win(g.h, sqrtf(fmaxf(0.f, 1.f - cr * cr))));
}

// one view's GBC contribution: value += w * detector(x), weight += w
template <typename Tp>
__device__ __forceinline__ void gbc_view(const BpCfg& c, const ViewMap& m, const Tp* __restrict__ img,
                                         const float x[3], float& val, float& wsum) {
    const float w = gbc_weight(c.gbc, m.S, x);
    wsum += w;
    if (w == 0.f) return;
    const float iw = __builtin_amdgcn_rcpf(m.W[0] * x[0] + m.W[1] * x[1] + m.W[2] * x[2] + m.w0);
    const float cu = (m.Au[0] * x[0] + m.Au[1] * x[1] + m.Au[2] * x[2] + m.au) * iw;
    const float cv = (m.Av[0] * x[0] + m.Av[1] * x[1] + m.Av[2] * x[2] + m.av) * iw;
    val += sample_det(c, m, img, cu, cv) * w;
}

// GBC backprojection (backproject_CB_voltex_GBC_cyl, CB_backproject.cu:1033-1108):
// the volume receives sum_views w * detector value, the secondary volume the
// sum of the weights w; sub-voxel means as in the other branches, both times
// the summand multiplier.  No mask, softening or Z powers apply.
template <typename Tv, typename Tp, typename Tm>
__global__ __launch_bounds__(kBz* kGbcRows) void k_backproject_gbc(
    BpCfg c, const ap_proj_spec* __restrict__ specs, uint32_t num_projs,
    Tv* __restrict__ vol, int64_t vs0, int64_t vs1, int64_t vs2,
    Tv* __restrict__ sec, int64_t ss0, int64_t ss1, int64_t ss2,
    const Tm* __restrict__ mult, int64_t ms0, int64_t ms1, int64_t ms2,
    const Tp* __restrict__ sino) {
    __shared__ ViewMap maps[kViewBatch];
    const uint32_t ck = blockIdx.x * kBz + threadIdx.x;
    const uint32_t cj = blockIdx.y * kGbcRows + threadIdx.y;
    const uint32_t ci = blockIdx.z;
    const bool active = ck < c.shp[2] && cj < c.shp[1] && ci < c.shp[0];
    const uint32_t gi = ci + c.off[0], gj = cj + c.off[1], gk = ck + c.off[2];
    const float xc[3] = {c.c000[0] + (float)gi * c.w, c.c000[1] + (float)gj * c.w,
                         c.c000[2] + (float)gk * c.w};
    const uint32_t nss = c.ssx * c.ssy * c.ssz;
    float val = 0.f, wsum = 0.f;
    const int tid = threadIdx.y * kBz + threadIdx.x;
    if (nss == 1) {
        for (uint32_t p0 = 0; p0 < num_projs; p0 += kViewBatch) {
            const uint32_t nb = min((uint32_t)kViewBatch, num_projs - p0);
            __syncthreads();
            if (tid < (int)nb) make_view_map(specs, p0 + tid, maps[tid]);
            __syncthreads();
            if (!active) continue;
            for (uint32_t q = 0; q < nb; q++)
                gbc_view(c, maps[q], sino + (int64_t)(p0 + q) * c.s0, xc, val, wsum);
        }
    } else if (active) {
        // sub-voxel loop order x, y, z (CB_backproject.cu:822-890)
        float vx = 0.f, wx = 0.f;
        for (uint32_t si = 0; si < c.ssx; si++) {
            float vy = 0.f, wy = 0.f;
            for (uint32_t sj = 0; sj < c.ssy; sj++) {
                float vz = 0.f, wz = 0.f;
                for (uint32_t sk = 0; sk < c.ssz; sk++) {
                    const float xs[3] = {xc[0] + (-0.5f + (0.5f + si) / c.ssx) * c.w,
                                         xc[1] + (-0.5f + (0.5f + sj) / c.ssy) * c.w,
                                         xc[2] + (-0.5f + (0.5f + sk) / c.ssz) * c.w};
                    float a = 0.f, b = 0.f;
                    for (uint32_t p = 0; p < num_projs; p++) {
                        ViewMap m;
                        make_view_map(specs, p, m);
                        gbc_view(c, m, sino + (int64_t)p * c.s0, xs, a, b);
                    }
                    vz += a;
                    wz += b;
                }
                vy += vz / c.ssz;
                wy += wz / c.ssz;
            }
            vx += vy / c.ssy;
            wx += wy / c.ssy;
        }
        val = vx / c.ssx;
        wsum = wx / c.ssx;
    }
    if (!active || (val == 0.f && wsum == 0.f)) return;
    float mval = 1.f;
    if (c.use_mult) mval = to_f32(mult[(int64_t)gi * ms0 + (int64_t)gj * ms1 + (int64_t)gk * ms2]);
    if (val != 0.f) {
        const int64_t o = (int64_t)gi * vs0 + (int64_t)gj * vs1 + (int64_t)gk * vs2;
        vol[o] = from_f32<Tv>(to_f32(vol[o]) + val * mval);
    }
    if (wsum != 0.f) {
        const int64_t o = (int64_t)gi * ss0 + (int64_t)gj * ss1 + (int64_t)gk * ss2;
        sec[o] = from_f32<Tv>(to_f32(sec[o]) + wsum * mval);
    }
}

// Integral of the GBC weight over a full cylindrical source locus for
// points (r, 0, 0), R = 1 (get_GBC_weight_integ, CB_backproject.cu:1111-1168):
// a workgroup sums kIntegAngles source angles x all heights for one r in
// FP64 and adds the sum, times the quadrature weight, to out[r].
constexpr int kIntegThreads = 256, kIntegAngles = 16;

__global__ __launch_bounds__(kIntegThreads) void k_gbc_weight_integ(GbcSoft g, float tan_va,
                                                                    const float* __restrict__ r_on_R, uint32_t n,
                                                                    double* __restrict__ out, uint32_t nz,
                                                                    uint32_t nt) {
    __shared__ double red[kIntegThreads / 64];
    const uint32_t ri = blockIdx.y;
    const float x[3] = {r_on_R[ri], 0.f, 0.f};
    const float maxz = tan_va * (1.f + r_on_R[n - 1]);
    const uint32_t a0 = blockIdx.x * kIntegAngles;
    const uint32_t na = min((uint32_t)kIntegAngles, nt - a0);
    double acc = 0.0;
    for (uint32_t e = threadIdx.x; e < na * nz; e += kIntegThreads) {
        const uint32_t a = a0 + e / nz, b = e % nz;
        const float t = (float)a / (float)nt * 3.1415926535898f;
        const float S[3] = {cosf(t), sinf(t), maxz * (float)b / (float)(nz - 1)};
        const float w = gbc_weight(g, S, x);
        acc += (double)(b == 0 ? 0.5f * w : w);   // trapezoid end weight at z = 0
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int w = 0; w < kIntegThreads / 64; w++) sum += red[w];
        const float q = (4.f * maxz * 3.1415926535898f) / ((float)nz * (float)nt);
        atomicAdd(&out[ri], sum * (double)q);
    }
}

}  // namespace

extern "C" int32_t ap_gbc_weight_integ(const float* r_on_R, uint32_t n, double* out, uint32_t num_z_samps,
                                       uint32_t num_ang_samps, const ap_config_cb_backproject* cfg,
                                       void* stream) {
    AP_REQUIRE(cfg, "NULL config");
    if (n == 0) return AP_OK;
    AP_REQUIRE(r_on_R && out, "NULL r_on_R or out");
    AP_REQUIRE(num_z_samps >= 2, "num_z_samps must be >= 2");
    AP_REQUIRE(num_ang_samps >= 1, "num_ang_samps must be >= 1");
    AP_REQUIRE(n <= 65535u, "at most 65535 radii per call");
    const GbcSoft g = make_gbc_soft(cfg, 1.0);
    const float tan_va = (float)std::tan((double)cfg->soft_vartheta_v);
    dim3 grid((num_ang_samps + kIntegAngles - 1) / kIntegAngles, n, 1);
    hipLaunchKernelGGL(k_gbc_weight_integ, grid, dim3(kIntegThreads), 0, (hipStream_t)stream, g, tan_va, r_on_R, n,
                       out, num_z_samps, num_ang_samps);
    return ap_check_launch("gbc_weight_integ");
}

int32_t launch_backproject_voxel(const ap_config_cb_backproject* cfg, const ap_proj_spec* specs,
                                 uint32_t num_projs, const ap_vol_tex_chunk_spec* chunk,
                                 const ap_array3* vol_desc, void* vol,
                                 const ap_array3* mult_desc, const void* multiplier,
                                 const ap_array3* sino_desc, const void* sino,
                                 const ap_array3* sec_desc, void* secondary, hipStream_t s) {
    BpCfg c;
    c.preset = cfg->backproj_weighting_preset;
    c.gbc = make_gbc_soft(cfg, cfg->Z_on_val_factor_val);
    c.exact = cfg->exact_line_integ_backproj != 0;
    c.all1s = cfg->backproject_all_1s != 0;
    c.mask = cfg->mask_to_cylinder != 0;
    c.mask_r2 = cfg->mask_to_cyl_rad * cfg->mask_to_cyl_rad;
    c.direct = cfg->approx_mode_direct_raycast != 0;
    c.bilinear = cfg->approx_mode_linear_interpolate != 0;
    c.use_mult = cfg->apply_summand_multiplier != 0 && multiplier != nullptr;
    c.sep = cfg->soften_edge_proportion;
    c.zl = cfg->Z_on_L_factor_power;
    c.zv = cfg->Z_on_val_factor_power;
    c.zval = cfg->Z_on_val_factor_val;
    c.inv_zval = (float)(1.0 / (double)cfg->Z_on_val_factor_val);
    c.ssx = cfg->ssx; c.ssy = cfg->ssy; c.ssz = cfg->ssz;
    c.w = chunk->voxwid;
    c.c000[0] = chunk->texpos.x + 0.5f * chunk->voxwid;
    c.c000[1] = chunk->texpos.y + 0.5f * chunk->voxwid;
    c.c000[2] = chunk->texpos.z + 0.5f * chunk->voxwid;
    c.off[0] = chunk->offsets.x; c.off[1] = chunk->offsets.y; c.off[2] = chunk->offsets.z;
    c.shp[0] = chunk->shape.x; c.shp[1] = chunk->shape.y; c.shp[2] = chunk->shape.z;
    c.nu = sino_desc->shape[1];
    c.nv = sino_desc->shape[2];
    c.s0 = sino_desc->strides[0]; c.s1 = sino_desc->strides[1]; c.s2 = sino_desc->strides[2];
    c.inv_sep = c.sep > 0.f ? 1.f / c.sep : 0.f;
    c.w3 = c.w * c.w * c.w;
    c.off32 = (double)(c.nu ? c.nu - 1 : 0) * (double)c.s1 + (double)(c.nv ? c.nv - 1 : 0) * (double)c.s2 <
              2147483647.0;
    c.pair = c.off32 && c.s2 == 1 && c.nv >= 2;
    c.grouped = !c.exact && (c.bilinear ? c.pair : c.off32) &&
                (double)c.s0 * 3.0 + (double)(c.nu ? c.nu - 1 : 0) * (double)c.s1 +
                        (double)(c.nv ? c.nv - 1 : 0) * (double)c.s2 < 2147483647.0;
    const int var = ap_tune_variant(AP_TUNE_BACKPROJECT);
    const uint32_t rows_y = (uint32_t)kBpRowsY[var];
    // four voxels per thread along z where a column is taller than two waves'
    // worth and the interpolating branch runs bp_col_views (which shares a
    // z-coherent view's u part between them)
    const uint32_t nss = c.ssx * c.ssy * c.ssz;
    const bool z4 = !c.exact && nss == 1 && (c.bilinear ? c.pair : c.off32) && c.shp[2] > 2 * (uint32_t)kBz;
    const uint32_t zper = z4 ? 4u * kBz : (uint32_t)kBz;
    dim3 block(kBz, rows_y, 1);
    dim3 grid((c.shp[2] + zper - 1) / zper, (c.shp[1] + rows_y - 1) / rows_y, c.shp[0]);
    const int32_t mdt = multiplier ? mult_desc->dtype : vol_desc->dtype;
    const int64_t* ms = multiplier ? mult_desc->strides : vol_desc->strides;
    return dispatch_dtype(vol_desc->dtype, [&](auto tv) -> int32_t {
        using Tv = decltype(tv);
        return dispatch_dtype(sino_desc->dtype, [&](auto tp) -> int32_t {
            using Tp = decltype(tp);
            return dispatch_dtype(mdt, [&](auto tm) -> int32_t {
                using Tm = decltype(tm);
                if constexpr (std::is_same<Tv, double>::value || std::is_same<Tp, double>::value ||
                              std::is_same<Tm, double>::value || std::is_same<Tp, bf16_t>::value) {
                    return ap_fail(AP_ERR_UNSUPPORTED, "unsupported dtype combination");
                } else {
                    if (c.preset == 7) {
                        dim3 gblock(kBz, kGbcRows, 1);
                        dim3 ggrid((c.shp[2] + kBz - 1) / kBz, (c.shp[1] + kGbcRows - 1) / kGbcRows, c.shp[0]);
                        hipLaunchKernelGGL((k_backproject_gbc<Tv, Tp, Tm>), ggrid, gblock, 0, s, c, specs,
                                           num_projs, (Tv*)vol, vol_desc->strides[0], vol_desc->strides[1],
                                           vol_desc->strides[2], (Tv*)secondary, sec_desc->strides[0],
                                           sec_desc->strides[1], sec_desc->strides[2], (const Tm*)multiplier,
                                           ms[0], ms[1], ms[2], (const Tp*)sino);
                        return ap_check_launch("backproject_gbc");
                    }
                    auto launch = [&](auto kern) {
                        hipLaunchKernelGGL(kern, grid, block, 0, s, c, specs, num_projs, (Tv*)vol,
                                           vol_desc->strides[0], vol_desc->strides[1], vol_desc->strides[2],
                                           (const Tm*)multiplier, ms[0], ms[1], ms[2], (const Tp*)sino);
                    };
                    if (z4) {
                        if (var == 0) launch(k_backproject_voxel<kBpRowsY[0], 4, Tv, Tp, Tm>);
                        if (var == 1) launch(k_backproject_voxel<kBpRowsY[1], 4, Tv, Tp, Tm>);
                        if (var == 2) launch(k_backproject_voxel<kBpRowsY[2], 4, Tv, Tp, Tm>);
                    } else {
                        if (var == 0) launch(k_backproject_voxel<kBpRowsY[0], 1, Tv, Tp, Tm>);
                        if (var == 1) launch(k_backproject_voxel<kBpRowsY[1], 1, Tv, Tp, Tm>);
                        if (var == 2) launch(k_backproject_voxel<kBpRowsY[2], 1, Tv, Tp, Tm>);
                    }
                    return ap_check_launch("backproject_voxel");
                }
            });
        });
    });
}

extern "C" int32_t ap_backproject(const ap_config_cb_backproject* cfg, const ap_proj_spec* specs,
                                  uint32_t num_projs, const ap_vol_tex_chunk_spec* chunk,
                                  const ap_array3* vol_desc, void* vol,
                                  const ap_array3* mult_desc, const void* multiplier,
                                  const ap_array3* sino_desc, const void* sino,
                                  const ap_array3* sec_desc, void* secondary, void* stream) {
    AP_REQUIRE(cfg && chunk && specs, "NULL config, chunk or specs");
    int32_t st = ap_check_array3(vol_desc, vol, "volume", true, true, false);
    if (st) return st;
    st = ap_check_array3(sino_desc, sino, "sino", true, false, false);
    if (st) return st;
    AP_REQUIRE(sino_desc->shape[0] == num_projs, "sino has %u views, num_projs=%u",
               sino_desc->shape[0], num_projs);
    if (cfg->backproj_weighting_preset == 7) {
        // the GBC kernel also writes the sum of its weights (CB_backproject.cu:987, 1072-1108)
        AP_REQUIRE(secondary != nullptr, "backproj_weighting_preset 7 needs a secondary volume");
        if ((st = ap_check_array3(sec_desc, secondary, "secondary volume", true, true, false))) return st;
        AP_REQUIRE(sec_desc->dtype == vol_desc->dtype, "secondary volume dtype != volume dtype");
        for (int a = 0; a < 3; a++)
            AP_REQUIRE(sec_desc->shape[a] == vol_desc->shape[a], "secondary volume shape != volume shape");
        AP_REQUIRE(secondary != vol, "the secondary volume must not be the volume");
        AP_REQUIRE(cfg->Z_on_val_factor_val != 0.f, "preset 7 needs the source radius in Z_on_val_factor_val");
    }
    AP_REQUIRE(cfg->ssx >= 1 && cfg->ssy >= 1 && cfg->ssz >= 1, "ssx/ssy/ssz must be >= 1");
    AP_REQUIRE(chunk->voxwid > 0.f, "voxwid must be positive");
    if ((cfg->apply_summand_multiplier != 0) != (multiplier != nullptr))
        return ap_fail(AP_ERR_INVALID,
                       "apply_summand_multiplier must be set if and only if a multiplier is given");
    if (multiplier) {
        if ((st = ap_check_array3(mult_desc, multiplier, "multiplier", true, true, false))) return st;
        for (int a = 0; a < 3; a++)
            AP_REQUIRE(mult_desc->shape[a] == vol_desc->shape[a], "multiplier shape != volume shape");
    }
    const uint32_t off[3] = {chunk->offsets.x, chunk->offsets.y, chunk->offsets.z};
    const uint32_t shp[3] = {chunk->shape.x, chunk->shape.y, chunk->shape.z};
    for (int a = 0; a < 3; a++)
        AP_REQUIRE((uint64_t)off[a] + shp[a] <= vol_desc->shape[a], "chunk exceeds the volume");
    if (num_projs == 0 || shp[0] == 0 || shp[1] == 0 || shp[2] == 0) return AP_OK;
    AP_REQUIRE(shp[0] <= 65535u, "chunk x extent above 65535");
    return launch_backproject_voxel(cfg, specs, num_projs, chunk, vol_desc, vol, mult_desc,
                                    multiplier, sino_desc, sino, sec_desc, secondary, (hipStream_t)stream);
}
