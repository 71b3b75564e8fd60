// forbild.hip -- analytic FORBILD primitives on the GPU: exact last-wins line
// integrals (ap_project_forbild), point-sampling voxeliser
// (ap_voxelise_forbild) and per-shape bounding boxes (ap_forbild_bboxes).
//
// Every primitive is reduced once per block to an affine map M (in FP64)
// sending it onto a reference shape -- the cube [-1,1]^3, the unit ball, the
// unit-radius cylinder |z| <= h, or the truncated cone |z| <= h,
// x^2 + y^2 <= (ca + cb z)^2 -- so a ray needs one 3x3 transform and one
// quadratic or slab solve per shape.  Ray-primitive parameters are FP64
// (grazing rays cancel catastrophically in FP32); the compositing sweep and
// accumulation are FP32/FP64 as the output dtype requires.
//
// Semantics: cuda/structs/cu/FORBILD_{Box,Sphere,Ellipsoid,Ellipsoid_free,
// Cylinder,Cylinder_x,Cylinder_y,Cylinder_z,Ellipt_Cyl,Ellipt_Cyl_x,
// Ellipt_Cyl_y,Ellipt_Cyl_z,Cone,Cone_x,Cone_y,Cone_z,ClipAxisSpec}.cu
// (intervals / membership; a cone's radius runs linearly from r1 at -l/2 to
// r2 at +l/2 along its axis),
// FORBILD_Shape.cu:474-540 (last shape wins), CB_project_texture_chunked.cu
// :421-593 (sub-pixels, multiply_before_adding, overflow counter).
#include "ap_common.h"

namespace {

enum PrimClass : int { kNone = 0, kCube = 1, kBall = 2, kCyl = 3, kCone = 4 };

struct Prim {
    double M[9];      // rows of the map r -> M (r - c)
    double c[3];
    double h;         // cylinder / cone half length
    double ca, cb;    // cone radius ca + cb z in the cone frame
    double cr;        // cone: max(|r1|, |r2|)
    float clip_ax[3][3];
    float clip_d[3];
    int clip_mode[3];
    float rho;
    int cls;
    float bc[3], br2;   // bounding sphere (centre, squared radius) for culling; br2 NaN: never culled
    float br;           // its radius
};

// Radius of a sphere about c holding the primitive: |M^-1 l| <= sigma_max(M^-1) |l|
// for the local point l, |l| <= lr (the local shape's radius), and sigma_max^2 =
// lambda_max((M^-1)^T M^-1) <= its largest absolute row sum (Gershgorin; exact
// when the frame is orthonormal, since the matrix is then diagonal).
__device__ void set_bound(Prim& P, double lr) {
    const double* m = P.M;
    const double c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8], c02 = m[3] * m[7] - m[4] * m[6];
    const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    const double id = 1.0 / det;
    // A = M^-1 = adj(M) / det
    const double A[9] = {c00 * id, (m[2] * m[7] - m[1] * m[8]) * id, (m[1] * m[5] - m[2] * m[4]) * id,
                         c01 * id, (m[0] * m[8] - m[2] * m[6]) * id, (m[2] * m[3] - m[0] * m[5]) * id,
                         c02 * id, (m[1] * m[6] - m[0] * m[7]) * id, (m[0] * m[4] - m[1] * m[3]) * id};
    double lam = 0.0;
    for (int i = 0; i < 3; i++) {
        double row = 0.0;
        for (int j = 0; j < 3; j++)
            row += fabs(A[i] * A[j] + A[3 + i] * A[3 + j] + A[6 + i] * A[6 + j]);
        lam = fmax(lam, row);
    }
    const double r = sqrt(lam) * lr * (1.0 + 1e-6);
    P.bc[0] = (float)P.c[0];
    P.bc[1] = (float)P.c[1];
    P.bc[2] = (float)P.c[2];
    P.br2 = (float)(r * r);   // inf or NaN (degenerate shape) never culls
    P.br = (float)r;
}

__device__ __forceinline__ d3 unit_of(ap_fvec3 a) {
    d3 v = fv(a);
    return v * (1.0 / sqrt(dot(v, v)));
}

__device__ void make_prim(const ap_forbild_shape& s, Prim& P) {
    d3 ex = mkd3(1, 0, 0), ey = mkd3(0, 1, 0), ez = mkd3(0, 0, 1);
    double sx = 1, sy = 1, sz = 1;
    P.h = 0;
    P.ca = P.cb = P.cr = 0;
    P.cls = kNone;
    switch (s.type) {
    case 1: P.cls = kCube; sx = 2.0 / s.dx; sy = 2.0 / s.dy; sz = 2.0 / s.dz; break;
    case 10: P.cls = kBall; sx = 1.0 / s.dx; sy = 1.0 / s.dy; sz = 1.0 / s.dz; break;
    case 11:
        P.cls = kBall;
        ex = unit_of(s.a_x); ey = unit_of(s.a_y); ez = unit_of(s.a_z);
        sx = 1.0 / s.dx; sy = 1.0 / s.dy; sz = 1.0 / s.dz;
        break;
    case 13: P.cls = kBall; sx = sy = sz = 1.0 / s.r; break;
    case 6: case 7: case 8: case 9: {
        P.cls = kCyl;
        d3 ax = s.type == 6 ? unit_of(s.axis)
              : s.type == 7 ? mkd3(1, 0, 0) : s.type == 8 ? mkd3(0, 1, 0) : mkd3(0, 0, 1);
        d3 t = fabs(ax.x) < 0.9 ? mkd3(1, 0, 0) : mkd3(0, 1, 0);
        d3 e1 = cross(ax, t);
        e1 = e1 * (1.0 / sqrt(dot(e1, e1)));
        ex = e1; ey = cross(ax, e1); ez = ax;
        sx = sy = 1.0 / s.r; sz = 1.0;
        P.h = 0.5 * (double)s.l;
        break;
    }
    case 2: case 3: case 4: case 5: {
        P.cls = kCone;
        d3 ax = s.type == 2 ? unit_of(s.axis)
              : s.type == 3 ? mkd3(1, 0, 0) : s.type == 4 ? mkd3(0, 1, 0) : mkd3(0, 0, 1);
        d3 t = fabs(ax.x) < 0.9 ? mkd3(1, 0, 0) : mkd3(0, 1, 0);
        d3 e1 = cross(ax, t);
        e1 = e1 * (1.0 / sqrt(dot(e1, e1)));
        ex = e1; ey = cross(ax, e1); ez = ax;
        P.h = 0.5 * (double)s.l;
        P.ca = 0.5 * ((double)s.r1 + (double)s.r2);
        P.cb = ((double)s.r2 - (double)s.r1) / (double)s.l;
        P.cr = fmax(fabs((double)s.r1), fabs((double)s.r2));
        break;
    }
    case 12: case 14: case 15: case 16:
        P.cls = kCyl;
        if (s.type == 12) { ex = unit_of(s.a_x); ey = unit_of(s.a_y); ez = unit_of(s.axis); }
        else if (s.type == 14) { ex = mkd3(0, 1, 0); ey = mkd3(0, 0, 1); ez = mkd3(1, 0, 0); }
        else if (s.type == 15) { ex = mkd3(0, 0, 1); ey = mkd3(1, 0, 0); ez = mkd3(0, 1, 0); }
        sx = 1.0 / s.dx; sy = 1.0 / s.dy; sz = 1.0;
        P.h = 0.5 * (double)s.l;
        break;
    default: break;
    }
    P.M[0] = ex.x * sx; P.M[1] = ex.y * sx; P.M[2] = ex.z * sx;
    P.M[3] = ey.x * sy; P.M[4] = ey.y * sy; P.M[5] = ey.z * sy;
    P.M[6] = ez.x * sz; P.M[7] = ez.y * sz; P.M[8] = ez.z * sz;
    P.c[0] = s.x; P.c[1] = s.y; P.c[2] = s.z;
    const ap_forbild_clip* cs[3] = {&s.cas1, &s.cas2, &s.cas3};
    for (int i = 0; i < 3; i++) {
        P.clip_ax[i][0] = cs[i]->axis.x;
        P.clip_ax[i][1] = cs[i]->axis.y;
        P.clip_ax[i][2] = cs[i]->axis.z;
        P.clip_d[i] = cs[i]->dotprod;
        P.clip_mode[i] = cs[i]->mode;
    }
    P.rho = s.rho;
    // local radius: unit cube corner, unit ball, unit disk x [-h, h], cone of radius cr x [-h, h]
    const double lr = P.cls == kCube ? sqrt(3.0)
                    : P.cls == kBall ? 1.0
                    : P.cls == kCyl  ? sqrt(1.0 + P.h * P.h)
                    : P.cls == kCone ? sqrt(P.cr * P.cr + P.h * P.h) : 0.0;
    set_bound(P, lr);
}

// 1 / x for a nonzero finite x: the hardware reciprocal refined by two Newton
// steps (error about 1e-16 relative; an IEEE division is ten instructions)
__device__ __forceinline__ double rcp64(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

// stable roots of A t^2 + 2 B t + C = 0; false if no real interval
__device__ __forceinline__ bool quad_roots(double A, double B, double C, double& t0, double& t1) {
    if (A == 0.0) return false;
    const double disc = B * B - A * C;
    if (!(disc > 0.0)) return false;
    const double sq = sqrt(disc);
    const double q = B > 0.0 ? -(B + sq) : -(B - sq);
    const double r1 = q * rcp64(A), r2 = q != 0.0 ? C * rcp64(q) : -r1;
    t0 = fmin(r1, r2);
    t1 = fmax(r1, r2);
    return true;
}

// Parametric interval (within [0,1]) of S + t d inside the primitive
__device__ bool prim_interval(const Prim& P, const d3& S, const d3& d, double& ta, double& tb) {
    ta = 0.0;
    tb = 1.0;
    for (int i = 0; i < 3; i++) {
        if (P.clip_mode[i] == 0) continue;
        const double ax = P.clip_ax[i][0], ay = P.clip_ax[i][1], az = P.clip_ax[i][2];
        double s0 = ax * S.x + ay * S.y + az * S.z - (double)P.clip_d[i];
        double sd = ax * d.x + ay * d.y + az * d.z;
        if (P.clip_mode[i] == 2) { s0 = -s0; sd = -sd; }   // keep s0 + t sd <= 0
        if (sd == 0.0) {
            if (s0 > 0.0) return false;
        } else {
            const double t = -s0 * rcp64(sd);
            if (sd > 0.0) tb = fmin(tb, t); else ta = fmax(ta, t);
        }
    }
    if (!(tb > ta)) return false;
    const double r0 = S.x - P.c[0], r1 = S.y - P.c[1], r2 = S.z - P.c[2];
    const double ls[3] = {P.M[0] * r0 + P.M[1] * r1 + P.M[2] * r2,
                          P.M[3] * r0 + P.M[4] * r1 + P.M[5] * r2,
                          P.M[6] * r0 + P.M[7] * r1 + P.M[8] * r2};
    const double ld[3] = {P.M[0] * d.x + P.M[1] * d.y + P.M[2] * d.z,
                          P.M[3] * d.x + P.M[4] * d.y + P.M[5] * d.z,
                          P.M[6] * d.x + P.M[7] * d.y + P.M[8] * d.z};
    double t0, t1;
    switch (P.cls) {
    case kCube:
        for (int a = 0; a < 3; a++) {
            if (ld[a] == 0.0) {
                if (fabs(ls[a]) > 1.0) return false;
                continue;
            }
            const double inv = rcp64(ld[a]);
            const double x0 = (-1.0 - ls[a]) * inv, x1 = (1.0 - ls[a]) * inv;
            ta = fmax(ta, fmin(x0, x1));
            tb = fmin(tb, fmax(x0, x1));
        }
        break;
    case kBall:
        if (!quad_roots(ld[0] * ld[0] + ld[1] * ld[1] + ld[2] * ld[2],
                        ls[0] * ld[0] + ls[1] * ld[1] + ls[2] * ld[2],
                        ls[0] * ls[0] + ls[1] * ls[1] + ls[2] * ls[2] - 1.0, t0, t1))
            return false;
        ta = fmax(ta, t0);
        tb = fmin(tb, t1);
        break;
    case kCyl: {
        if (ld[2] == 0.0) {
            if (fabs(ls[2]) > P.h) return false;
        } else {
            const double inv = rcp64(ld[2]);
            const double x0 = (-P.h - ls[2]) * inv, x1 = (P.h - ls[2]) * inv;
            ta = fmax(ta, fmin(x0, x1));
            tb = fmin(tb, fmax(x0, x1));
        }
        const double A = ld[0] * ld[0] + ld[1] * ld[1];
        const double C = ls[0] * ls[0] + ls[1] * ls[1] - 1.0;
        if (A == 0.0) {
            if (C > 0.0) return false;
        } else {
            if (!quad_roots(A, ls[0] * ld[0] + ls[1] * ld[1], C, t0, t1)) return false;
            ta = fmax(ta, t0);
            tb = fmin(tb, t1);
        }
        break;
    }
    case kCone: {
        // axial slab |z| <= h
        if (ld[2] == 0.0) {
            if (fabs(ls[2]) > P.h) return false;
        } else {
            const double inv = rcp64(ld[2]);
            const double x0 = (-P.h - ls[2]) * inv, x1 = (P.h - ls[2]) * inv;
            ta = fmax(ta, fmin(x0, x1));
            tb = fmin(tb, fmax(x0, x1));
        }
        if (!(tb > ta)) return false;
        // f(t) = x^2 + y^2 - (ca + cb z)^2 = A t^2 + 2 B t + C.  Inside the slab
        // only the nappe with ca + cb z >= 0 exists and the truncated cone is
        // convex, so {f <= 0} meets [ta, tb] in one interval.
        const double g0 = P.ca + P.cb * ls[2], gd = P.cb * ld[2];
        const double A = ld[0] * ld[0] + ld[1] * ld[1] - gd * gd;
        const double B = ls[0] * ld[0] + ls[1] * ld[1] - g0 * gd;
        const double C = ls[0] * ls[0] + ls[1] * ls[1] - g0 * g0;
        if (A == 0.0) {
            if (B == 0.0) {
                if (C > 0.0) return false;
            } else if (B > 0.0) {
                tb = fmin(tb, -C * rcp64(2.0 * B));
            } else {
                ta = fmax(ta, -C * rcp64(2.0 * B));
            }
        } else if (A > 0.0) {
            if (!quad_roots(A, B, C, t0, t1)) return false;
            ta = fmax(ta, t0);
            tb = fmin(tb, t1);
        } else if (quad_roots(A, B, C, t0, t1)) {
            // f <= 0 outside (t0, t1): keep the piece on the real nappe
            const double la = ta, lb = fmin(tb, t0), ra = fmax(ta, t1), rb = tb;
            const bool left = lb > la, right = rb > ra;
            if (left && right) {
                const double zl = ls[2] + 0.5 * (la + lb) * ld[2];
                if (P.ca + P.cb * zl >= 0.0) tb = lb;
                else ta = ra;
            } else if (left) {
                tb = lb;
            } else if (right) {
                ta = ra;
            } else {
                return false;
            }
        }   // A < 0 without real roots: f < 0 on the whole line
        break;
    }
    default:
        return false;
    }
    return tb > ta;
}

__device__ bool prim_contains(const Prim& P, const d3& p) {
    for (int i = 0; i < 3; i++) {
        if (P.clip_mode[i] == 0) continue;
        const double dp = (double)P.clip_ax[i][0] * p.x + (double)P.clip_ax[i][1] * p.y +
                          (double)P.clip_ax[i][2] * p.z;
        if (P.clip_mode[i] == 1 && !(dp <= (double)P.clip_d[i])) return false;
        if (P.clip_mode[i] == 2 && !(dp > (double)P.clip_d[i])) return false;
    }
    const double r0 = p.x - P.c[0], r1 = p.y - P.c[1], r2 = p.z - P.c[2];
    const double a = P.M[0] * r0 + P.M[1] * r1 + P.M[2] * r2;
    const double b = P.M[3] * r0 + P.M[4] * r1 + P.M[5] * r2;
    const double e = P.M[6] * r0 + P.M[7] * r1 + P.M[8] * r2;
    switch (P.cls) {
    case kCube: return fabs(a) <= 1.0 && fabs(b) <= 1.0 && fabs(e) <= 1.0;
    case kBall: return a * a + b * b + e * e <= 1.0;
    case kCyl: return fabs(e) <= P.h && a * a + b * b <= 1.0;
    case kCone: {
        const double g = P.ca + P.cb * e;
        return fabs(e) <= P.h && a * a + b * b <= g * g;
    }
    default: return false;
    }
}

constexpr int kProjThreads = 256;  // 4 waves share the staged shapes
constexpr int kShapeBatch = 64;    // shapes staged in LDS at a time
constexpr int kUnion = 8;          // disjoint covered pieces kept per ray before the exact fallback
constexpr uint32_t kRefMaxIntersections = 1000;   // reference's overflow threshold

// Last-wins compositing (FORBILD_Shape.cu:474-540): the density at a point is
// that of the highest-index shape containing it, so the integral is
//   sum_q rho_q |I_q \ (union of I_q' for q' > q)|.
// Shapes are visited from the last to the first; the union of the intervals
// seen so far is kept as at most kUnion disjoint pieces in registers (empty
// pieces are [3, 2]).  Adding an interval I: its covered length is the sum of
// its overlaps with the pieces; then every piece it overlaps or touches is
// absorbed into it (one pass suffices: the pieces are disjoint, so a piece
// that meets the grown interval meets I itself) and the result takes a free
// piece.  Returns false when no piece is free.
struct Union {
    double a[kUnion], b[kUnion];
};

__device__ __forceinline__ void union_clear(Union& U) {
#pragma unroll
    for (int i = 0; i < kUnion; i++) {
        U.a[i] = 3.0;
        U.b[i] = 2.0;
    }
}

__device__ __forceinline__ bool union_add(Union& U, double ta, double tb, double& uncovered) {
    double cov = 0.0, na = ta, nb = tb;
#pragma unroll
    for (int i = 0; i < kUnion; i++) {
        cov += fmax(0.0, fmin(tb, U.b[i]) - fmax(ta, U.a[i]));
        if (U.a[i] <= nb && U.b[i] >= na) {
            na = fmin(na, U.a[i]);
            nb = fmax(nb, U.b[i]);
            U.a[i] = 3.0;
            U.b[i] = 2.0;
        }
    }
    uncovered = (tb - ta) - cov;
    bool placed = false;
#pragma unroll
    for (int i = 0; i < kUnion; i++) {
        if (!placed && U.a[i] > U.b[i]) {
            U.a[i] = na;
            U.b[i] = nb;
            placed = true;
        }
    }
    return placed;
}

// Exact last-wins integral recomputed from global memory at every event; for
// rays whose covered set has more than kUnion pieces
__device__ double last_wins_exact(const ap_forbild_shape* __restrict__ shapes, uint32_t K, const d3& S,
                                  const d3& d) {
    double val = 0.0, prog = -1.0;
    for (uint32_t guard = 0; guard < 4 * K + 4; guard++) {
        int top = -1;
        double topb = 0.0, nxt = 2.0;
        float toprho = 0.f;
        for (int q = (int)K - 1; q >= 0; q--) {
            Prim P;
            make_prim(shapes[q], P);
            double ta, tb;
            if (!prim_interval(P, S, d, ta, tb)) continue;
            if (top < 0 && ta <= prog && prog < tb) {
                top = q; topb = tb; toprho = P.rho;
            } else if (top < 0 && ta > prog && ta < nxt) {
                nxt = ta;   // candidate next start (no cover yet)
            }
        }
        if (top >= 0) {
            double e = topb;
            for (int q = (int)K - 1; q > top; q--) {
                Prim P;
                make_prim(shapes[q], P);
                double ta, tb;
                if (prim_interval(P, S, d, ta, tb) && ta > prog && ta < e) e = ta;
            }
            val += (e - prog) * (double)toprho;
            prog = e;
        } else {
            if (nxt >= 2.0) break;
            prog = nxt;
        }
    }
    return val;
}

// The FP64 intersection chains are latency that more waves hide, even at the
// cost of spilling part of the union state to scratch.  kOne: one ray per
// pixel (ssu = ssv = 1), so the view's detector frame is dead once the ray is
// formed: 7 waves per SIMD.  With sub-rays the frame stays live for the next
// one: 4 waves.  40 primitives x 720 x 512^2: 22.9 ms at 3 waves (both
// cases); kOne 18.2 ms at 4 waves (no spills), 16.3 at 6, 15.6 at 7, 22.8 at 8.
template <typename Tp, bool kOne>
__global__ __launch_bounds__(kProjThreads, kOne ? 7 : 4) void k_project_forbild(
    float mba, int additive, uint32_t ssu, uint32_t ssv, const ap_proj_spec* __restrict__ specs,
    const ap_forbild_shape* __restrict__ shapes, uint32_t K, uint32_t nu, uint32_t nv,
    int64_t s0, int64_t s1, int64_t s2, Tp* __restrict__ sino, uint32_t* __restrict__ oom) {
    __shared__ Prim prims[kShapeBatch];
    const uint32_t pix = blockIdx.x * kProjThreads + threadIdx.x;
    const uint32_t p = blockIdx.y;
    const bool in = pix < nu * nv;
    const uint32_t j = in ? pix / nv : 0, k = in ? pix % nv : 0;
    const ViewGeom g = load_view(specs, p);
    double tot = 0.0;
    bool overflow = false;
    if constexpr (kOne) ssu = ssv = 1;
    for (uint32_t su = 0; su < ssu; su++) {
        double accu = 0.0;
        for (uint32_t sv = 0; sv < ssv; sv++) {
            const double ju = (double)j - 0.5 + ((double)su + 0.5) / ssu;
            const double kv = (double)k - 0.5 + ((double)sv + 0.5) / ssv;
            const double a = ju - 0.5 * (double)(g.nu - 1), b = kv - 0.5 * (double)(g.nv - 1);
            const d3 Q = g.D + g.U * a + g.V * b;
            const d3 d = Q - g.S;
            const double len = sqrt(dot(d, d));
            const bool live = in && len > 0.0;
            // the ray in FP32 for the bounding-sphere cull
            const float Sf[3] = {(float)g.S.x, (float)g.S.y, (float)g.S.z};
            const float df[3] = {(float)d.x, (float)d.y, (float)d.z};
            const float idd = 1.f / (df[0] * df[0] + df[1] * df[1] + df[2] * df[2]);
            uint32_t nhit = 0;
            bool spill = false;
            double val = 0.0;
            Union U;
            union_clear(U);
            // batches from the last shape down, each visited last to first
            for (uint32_t end = K; end > 0;) {
                const uint32_t nb = min((uint32_t)kShapeBatch, end);
                const uint32_t k0 = end - nb;
                __syncthreads();
                if (threadIdx.x < nb) make_prim(shapes[k0 + threadIdx.x], prims[threadIdx.x]);
                __syncthreads();
                end = k0;
                if (!live || spill) continue;
                for (int q = (int)nb - 1; q >= 0; q--) {
                    // the line misses the bounding sphere: squared distance of
                    // its centre from the line, with a margin of 1e-6 |v|^2 for
                    // the FP32 cancellation (a wave skips the shape when all
                    // its rays miss)
                    const Prim& P = prims[q];
                    const float vx = P.bc[0] - Sf[0], vy = P.bc[1] - Sf[1], vz = P.bc[2] - Sf[2];
                    const float vd = vx * df[0] + vy * df[1] + vz * df[2];
                    const float vv = vx * vx + vy * vy + vz * vz;
                    if (vv - vd * vd * idd > P.br2 + 1e-6f * vv) continue;
                    double ta, tb;
                    if (!prim_interval(P, g.S, d, ta, tb)) continue;
                    // more than the reference's limit of intersections: flag
                    // it and ignore the rest (the reference drops the
                    // highest-index ones, this loop the lowest)
                    if (++nhit > kRefMaxIntersections) { overflow = true; continue; }
                    const double rho = (double)prims[q].rho;
                    if (additive) { val += (tb - ta) * rho; continue; }
                    double un;
                    if (!union_add(U, ta, tb, un)) { spill = true; break; }
                    val += un * rho;
                }
            }
            if (live && spill) val = last_wins_exact(shapes, K, g.S, d);
            accu += val * len;
        }
        tot += accu / ssv;
    }
    tot /= ssu;
    if (!in) return;
    if (overflow && oom) atomicAdd(oom, 1u);
    const int64_t o = (int64_t)p * s0 + (int64_t)j * s1 + (int64_t)k * s2;
    sino[o] = (Tp)((double)sino[o] + (double)mba * tot);
}

template <typename Tv>
__global__ __launch_bounds__(256) void k_voxelise(
    const ap_forbild_shape* __restrict__ shapes, uint32_t K, double vw, double tpx, double tpy,
    double tpz, uint32_t ox, uint32_t oy, uint32_t oz, uint32_t cx, uint32_t cy, uint32_t cz,
    Tv* __restrict__ vol, int64_t v0, int64_t v1, int64_t v2, uint32_t ss) {
    __shared__ Prim prims[kShapeBatch];
    const uint32_t k = blockIdx.x * 64 + threadIdx.x;
    const uint32_t j = blockIdx.y * 4 + threadIdx.y;
    const uint32_t i = blockIdx.z;
    const bool in = k < cz && j < cy && i < cx;
    const uint32_t gi = i + ox, gj = j + oy, gk = k + oz;
    const int tid = threadIdx.y * 64 + threadIdx.x;
    const uint32_t npt = ss * ss * ss;
    // per sub-point value of the last containing shape; ss <= 4 => <= 64 points
    float val[64];
    for (uint32_t q = 0; q < npt && q < 64; q++) val[q] = 0.f;
    float val1 = 0.f;   // the one point of ss == 1, kept in a register
    // voxel centre in FP32 for the bounding-sphere cull; every sub-point lies
    // within the voxel's half-diagonal of it
    const float cf[3] = {(float)(tpx + (gi + 0.5) * vw), (float)(tpy + (gj + 0.5) * vw),
                         (float)(tpz + (gk + 0.5) * vw)};
    const float reach = (float)(0.8660254037844386 * vw);
    for (uint32_t k0 = 0; k0 < K; k0 += kShapeBatch) {
        const uint32_t nb = min((uint32_t)kShapeBatch, K - k0);
        __syncthreads();
        if (tid < (int)nb) make_prim(shapes[k0 + tid], prims[tid]);
        __syncthreads();
        if (!in) continue;
        for (uint32_t s = 0; s < nb; s++) {
            const Prim& P = prims[s];
            // outside the bounding sphere grown by the half-diagonal, with a
            // margin for the FP32 centre (shapes are visited in order, so the
            // last containing shape still wins)
            const float dx = cf[0] - P.bc[0], dy = cf[1] - P.bc[1], dz = cf[2] - P.bc[2];
            const float rr = P.br * 1.0001f + reach + 1e-3f;
            if (dx * dx + dy * dy + dz * dz > rr * rr) continue;
            if (ss == 1) {
                if (prim_contains(P, mkd3(tpx + (gi + 0.5) * vw, tpy + (gj + 0.5) * vw, tpz + (gk + 0.5) * vw)))
                    val1 = P.rho;
                continue;
            }
            for (uint32_t a = 0, q = 0; a < ss; a++)
                for (uint32_t b = 0; b < ss; b++)
                    for (uint32_t e = 0; e < ss; e++, q++) {
                        const double fx = ((double)a - 0.5 * (ss - 1.0)) / ss;
                        const double fy = ((double)b - 0.5 * (ss - 1.0)) / ss;
                        const double fz = ((double)e - 0.5 * (ss - 1.0)) / ss;
                        const d3 pnt = mkd3(tpx + (gi + 0.5 + fx) * vw, tpy + (gj + 0.5 + fy) * vw,
                                            tpz + (gk + 0.5 + fz) * vw);
                        if (prim_contains(P, pnt)) val[q] = P.rho;
                    }
        }
    }
    if (!in) return;
    float tot = val1;
    if (ss != 1) {
        tot = 0.f;
        for (uint32_t q = 0; q < npt; q++) tot += val[q];
        tot /= (float)npt;
    }
    const int64_t o = (int64_t)gi * v0 + (int64_t)gj * v1 + (int64_t)gk * v2;
    vol[o] = from_f32<Tv>(to_f32(vol[o]) + tot);
}

__global__ void k_bboxes(const ap_forbild_shape* __restrict__ shapes, uint32_t K, float* out) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= K) return;
    Prim P;
    make_prim(shapes[q], P);
    // extent of the reference shape mapped back: r = c + M^-1 y.  For the
    // unit reference shapes, half-extent along world axis w is
    //   cube: sum_a |Minv[w][a]|, ball: |Minv row w|_2,
    //   cylinder: |(Minv[w][0],Minv[w][1])|_2 + h |Minv[w][2]|.
    const double* M = P.M;
    const double det = M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
                       M[2] * (M[3] * M[7] - M[4] * M[6]);
    double Mi[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (det != 0.0) {
        const double id = 1.0 / det;
        Mi[0] = (M[4] * M[8] - M[5] * M[7]) * id;
        Mi[1] = (M[2] * M[7] - M[1] * M[8]) * id;
        Mi[2] = (M[1] * M[5] - M[2] * M[4]) * id;
        Mi[3] = (M[5] * M[6] - M[3] * M[8]) * id;
        Mi[4] = (M[0] * M[8] - M[2] * M[6]) * id;
        Mi[5] = (M[2] * M[3] - M[0] * M[5]) * id;
        Mi[6] = (M[3] * M[7] - M[4] * M[6]) * id;
        Mi[7] = (M[1] * M[6] - M[0] * M[7]) * id;
        Mi[8] = (M[0] * M[4] - M[1] * M[3]) * id;
    }
    for (int w = 0; w < 3; w++) {
        const double a = Mi[3 * w], b = Mi[3 * w + 1], e = Mi[3 * w + 2];
        double h;
        if (P.cls == kCube) h = fabs(a) + fabs(b) + fabs(e);
        else if (P.cls == kBall) h = sqrt(a * a + b * b + e * e);
        else if (P.cls == kCyl) h = sqrt(a * a + b * b) + P.h * fabs(e);
        else if (P.cls == kCone) h = P.cr * sqrt(a * a + b * b) + P.h * fabs(e);
        else h = 0.0;
        out[6 * q + w] = (float)(P.c[w] - h);
        out[6 * q + 3 + w] = (float)(P.c[w] + h);
    }
}

int32_t check_shapes(const ap_forbild_shape* shapes, uint32_t K) {
    if (K > 0 && !shapes) return ap_fail(AP_ERR_INVALID, "shapes is NULL");
    return AP_OK;
}

}  // namespace

extern "C" int32_t ap_project_forbild(const ap_config_cb_project_forbild* cfg,
                                      const ap_proj_spec* specs, uint32_t num_projs,
                                      const ap_forbild_shape* shapes, uint32_t num_shapes,
                                      const ap_array3* sino_desc, void* sino,
                                      uint32_t* out_of_mem_count, void* stream) {
    AP_REQUIRE(cfg && specs, "NULL config or specs");
    AP_REQUIRE(cfg->ssu >= 1 && cfg->ssv >= 1, "ssu and ssv must be >= 1");
    int32_t st = check_shapes(shapes, num_shapes);
    if (st) return st;
    if ((st = ap_check_array3(sino_desc, sino, "sino", false, false, true))) return st;
    AP_REQUIRE(sino_desc->shape[0] == num_projs, "sino has %u views, num_projs=%u",
               sino_desc->shape[0], num_projs);
    const uint32_t nu = sino_desc->shape[1], nv = sino_desc->shape[2];
    if (num_projs == 0 || nu == 0 || nv == 0) return AP_OK;
    AP_REQUIRE(num_projs <= 65535u, "at most 65535 views per call");
    const uint64_t npix = (uint64_t)nu * nv;
    AP_REQUIRE(npix < (1ull << 31), "detector too large");
    dim3 grid((uint32_t)((npix + kProjThreads - 1) / kProjThreads), num_projs, 1);
    hipStream_t s = (hipStream_t)stream;
    const bool one = cfg->ssu == 1 && cfg->ssv == 1;
    if (sino_desc->dtype == AP_F32) {
        auto kern = one ? k_project_forbild<float, true> : k_project_forbild<float, false>;
        hipLaunchKernelGGL(kern, grid, dim3(kProjThreads), 0, s,
                           cfg->multiply_before_adding, (int)cfg->overlap_addition, cfg->ssu, cfg->ssv,
                           specs, shapes, num_shapes, nu, nv, sino_desc->strides[0],
                           sino_desc->strides[1], sino_desc->strides[2], (float*)sino, out_of_mem_count);
    } else if (sino_desc->dtype == AP_F64) {
        auto kern = one ? k_project_forbild<double, true> : k_project_forbild<double, false>;
        hipLaunchKernelGGL(kern, grid, dim3(kProjThreads), 0, s,
                           cfg->multiply_before_adding, (int)cfg->overlap_addition, cfg->ssu, cfg->ssv,
                           specs, shapes, num_shapes, nu, nv, sino_desc->strides[0],
                           sino_desc->strides[1], sino_desc->strides[2], (double*)sino, out_of_mem_count);
    } else {
        return ap_fail(AP_ERR_UNSUPPORTED, "FORBILD sinogram must be F32 or F64");
    }
    return ap_check_launch("project_forbild");
}

extern "C" int32_t ap_voxelise_forbild(const ap_forbild_shape* shapes, uint32_t num_shapes,
                                       const ap_vol_tex_chunk_spec* chunk,
                                       const ap_array3* vol_desc, void* vol,
                                       uint32_t supersample, void* stream) {
    AP_REQUIRE(chunk, "chunk is NULL");
    AP_REQUIRE(supersample >= 1 && supersample <= 4, "supersample must be in 1..4");
    AP_REQUIRE(chunk->voxwid > 0.f, "voxwid must be positive");
    int32_t st = check_shapes(shapes, num_shapes);
    if (st) return st;
    if ((st = ap_check_array3(vol_desc, vol, "volume", true, true, false))) return st;
    const uint32_t off[3] = {chunk->offsets.x, chunk->offsets.y, chunk->offsets.z};
    const uint32_t shp[3] = {chunk->shape.x, chunk->shape.y, chunk->shape.z};
    for (int a = 0; a < 3; a++)
        AP_REQUIRE((uint64_t)off[a] + shp[a] <= vol_desc->shape[a], "chunk exceeds the volume");
    if (shp[0] == 0 || shp[1] == 0 || shp[2] == 0) return AP_OK;
    AP_REQUIRE(shp[0] <= 65535u, "chunk x extent above 65535");
    dim3 grid((shp[2] + 63) / 64, (shp[1] + 3) / 4, shp[0]);
    hipStream_t s = (hipStream_t)stream;
    return dispatch_dtype(vol_desc->dtype, [&](auto tv) -> int32_t {
        using Tv = decltype(tv);
        if constexpr (std::is_same<Tv, double>::value) {
            return ap_fail(AP_ERR_UNSUPPORTED, "float64 volumes are not supported");
        } else {
            hipLaunchKernelGGL(k_voxelise<Tv>, grid, dim3(64, 4, 1), 0, s, shapes, num_shapes,
                               (double)chunk->voxwid, (double)chunk->texpos.x, (double)chunk->texpos.y,
                               (double)chunk->texpos.z, off[0], off[1], off[2], shp[0], shp[1], shp[2],
                               (Tv*)vol, vol_desc->strides[0], vol_desc->strides[1],
                               vol_desc->strides[2], supersample);
            return ap_check_launch("voxelise_forbild");
        }
    });
}

extern "C" int32_t ap_forbild_bboxes(const ap_forbild_shape* shapes, uint32_t num_shapes,
                                     float* bboxes, void* stream) {
    int32_t st = check_shapes(shapes, num_shapes);
    if (st) return st;
    AP_REQUIRE(num_shapes == 0 || bboxes, "bboxes is NULL");
    if (num_shapes == 0) return AP_OK;
    hipLaunchKernelGGL(k_bboxes, dim3((num_shapes + 127) / 128), dim3(128), 0, (hipStream_t)stream,
                       shapes, num_shapes, bboxes);
    return ap_check_launch("forbild_bboxes");
}
