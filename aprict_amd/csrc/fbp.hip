// fbp.hip -- the per-projection steps of the GBC / FBP filter around the row
// filter (FBPFilterer.filter_bulk_axis1_stream, cb/ramp_filter.py:577-886):
// zero padding to the expanded detector, rotation by -zeta (bilinear, zeros
// outside), the cone or cosine weight and the letterbox before the filter
// (ap_fbp_rotate_weight); rotation by +zeta and cropping after it
// (ap_fbp_unrotate).
//
// One thread per output pixel, 64 lanes along v (coalesced) x 4 rows along
// u; one workgroup row of blocks per view.  Thread 0 of each workgroup
// derives the view's rotation and letterbox into LDS.  Coordinates are
// mapped in FP64 (a pixel index near 1000 would lose 6e-5 pixel in FP32);
// the interpolation weights and the data are FP32.
#include "ap_common.h"

namespace {

constexpr int kLanesV = 64, kRowsU = 4;

struct RotView {
    double c, s;       // out(o) = in(M (o - centre) + centre), M = [[c, s], [-s, c]]
    bool copy;         // angle 0: exact copy
    int k1, k2;        // letterbox columns: zero outside [k1, k2]
    float m1, m2;      // weights of columns k1 and k2
};

// Letterbox of FBPFilterer.get_letterbox_slice and its application
// (cb/ramp_filter.py:470-500, 689-754): the v range m +- d of the rotated
// detector, m = (Ev - 1) / 2, d = sqrt(max(cos^2 z / cos^2 zmax - 1, 0)) / pw,
// with the end columns rounded half to even (Python's round) and weighted by
// the fraction of them inside the range.
__device__ void letterbox(float z, double zmax, double pw, uint32_t Ev, RotView& r) {
    const double cz = (double)(float)cos((double)z);   // the reference takes cos of the FP32 angle in FP32
    const double q = cz / cos(zmax);
    const double disc = (1.0 / pw) * sqrt(fmax(q * q - 1.0, 0.0));
    const double mid = 0.5 * ((double)Ev - 1.0);
    const double start = mid - disc, stop = mid + disc;
    r.k1 = (int)rint(start);
    r.k2 = (int)rint(stop);
    if (r.k1 == r.k2) {
        r.m1 = r.m2 = (float)(stop - start);
    } else {
        r.m1 = (float)(((double)r.k1 + 0.5) - start);
        r.m2 = (float)(stop - ((double)r.k2 - 0.5));
    }
}

// bilinear sample of a [nu][nv] image at (xu, xv), taps outside read 0
__device__ __forceinline__ float bilinear0(const float* __restrict__ img, int64_t s1, int64_t s2, int nu, int nv,
                                           double xu, double xv) {
    const double fu = floor(xu), fv = floor(xv);
    if (!(fu >= -1.0 && fv >= -1.0 && fu < (double)nu && fv < (double)nv)) return 0.f;   // also NaN
    const int a = (int)fu, b = (int)fv;
    const float tu = (float)(xu - fu), tv = (float)(xv - fv);
    auto rd = [&](int i, int k) -> float {
        return (i >= 0 && k >= 0 && i < nu && k < nv) ? img[(int64_t)i * s1 + (int64_t)k * s2] : 0.f;
    };
    const float v00 = rd(a, b), v10 = rd(a + 1, b), v01 = rd(a, b + 1), v11 = rd(a + 1, b + 1);
    return (1.f - tu) * ((1.f - tv) * v00 + tv * v01) + tu * ((1.f - tv) * v10 + tv * v11);
}

__global__ __launch_bounds__(kLanesV* kRowsU) void k_fbp_rotate_weight(
    uint32_t nu, uint32_t nv, uint32_t Eu, uint32_t Ev, const float* __restrict__ in, int64_t i0, int64_t i1,
    int64_t i2, float* __restrict__ out, int64_t o0, int64_t o1, int64_t o2, const float* __restrict__ angles,
    double pw, double zmax, int mode) {
    __shared__ RotView rv;
    const uint32_t p = blockIdx.z;
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const float z = angles ? angles[p] : 0.f;
        rv.copy = z == 0.f;
        double sn, cs;
        sincos((double)z, &sn, &cs);
        rv.c = cs;
        rv.s = -sn;   // rotation by -z
        rv.k1 = 0;
        rv.k2 = (int)Ev - 1;
        rv.m1 = rv.m2 = 1.f;
        if (mode == AP_FBP_GBC) letterbox(z, zmax, pw, Ev, rv);
    }
    __syncthreads();
    const uint32_t iv = blockIdx.x * kLanesV + threadIdx.x;
    const uint32_t iu = blockIdx.y * kRowsU + threadIdx.y;
    if (iu >= Eu || iv >= Ev) return;
    const int pu = (int)(Eu - nu) / 2, pv = (int)(Ev - nv) / 2;
    const float* img = in + (int64_t)p * i0;
    float val;
    if (rv.copy) {
        const int a = (int)iu - pu, b = (int)iv - pv;
        val = (a >= 0 && b >= 0 && a < (int)nu && b < (int)nv) ? img[(int64_t)a * i1 + (int64_t)b * i2] : 0.f;
    } else {
        const double cu = 0.5 * ((double)Eu - 1.0), cv = 0.5 * ((double)Ev - 1.0);
        const double du = (double)iu - cu, dv = (double)iv - cv;
        const double xu = rv.c * du + rv.s * dv + cu - (double)pu;
        const double xv = -rv.s * du + rv.c * dv + cv - (double)pv;
        val = bilinear0(img, i1, i2, (int)nu, (int)nv, xu, xv);
    }
    // detector weight on the plane at unit distance (cb/ramp_filter.py:410-449),
    // coordinates rounded to FP32 as the reference's FP32 grids are
    const float pw32 = (float)pw;
    const float uu = pw32 * ((float)iu - 0.5f * (float)(Eu - 1));
    const float vv = pw32 * ((float)iv - 0.5f * (float)(Ev - 1));
    const float r2 = (1.f + vv * vv) + uu * uu;
    if (mode == AP_FBP_FDK) {
        val *= 1.f / sqrtf(r2);
    } else {
        const double a = (double)sqrtf(1.f + vv * vv);
        val = (float)((double)val / (a * a * a * (double)sqrtf(r2)));
    }
    if (mode == AP_FBP_GBC) {
        const int k = (int)iv;
        if (k < rv.k1 || k > rv.k2) val = 0.f;
        else if (rv.k1 == rv.k2) val *= rv.m1;
        else if (k == rv.k1) val *= rv.m1;
        else if (k == rv.k2) val *= rv.m2;
    }
    out[(int64_t)p * o0 + (int64_t)iu * o1 + (int64_t)iv * o2] = val;
}

__global__ __launch_bounds__(kLanesV* kRowsU) void k_fbp_unrotate(
    uint32_t nu, uint32_t nv, uint32_t Eu, uint32_t Ev, const float* __restrict__ in, int64_t i0, int64_t i1,
    int64_t i2, float* __restrict__ out, int64_t o0, int64_t o1, int64_t o2, const float* __restrict__ angles) {
    __shared__ double cs_sn[2];
    __shared__ bool copy;
    const uint32_t p = blockIdx.z;
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const float z = angles ? angles[p] : 0.f;
        copy = z == 0.f;
        sincos((double)z, &cs_sn[1], &cs_sn[0]);   // rotation by +z
    }
    __syncthreads();
    const uint32_t jv = blockIdx.x * kLanesV + threadIdx.x;
    const uint32_t ju = blockIdx.y * kRowsU + threadIdx.y;
    if (ju >= nu || jv >= nv) return;
    const int pu = (int)(Eu - nu) / 2, pv = (int)(Ev - nv) / 2;
    const float* img = in + (int64_t)p * i0;
    float val;
    if (copy) {
        val = img[(int64_t)(ju + pu) * i1 + (int64_t)(jv + pv) * i2];
    } else {
        const double c = cs_sn[0], s = cs_sn[1];
        const double cu = 0.5 * ((double)Eu - 1.0), cv = 0.5 * ((double)Ev - 1.0);
        const double du = (double)(ju + pu) - cu, dv = (double)(jv + pv) - cv;
        val = bilinear0(img, i1, i2, (int)Eu, (int)Ev, c * du + s * dv + cu, -s * du + c * dv + cv);
    }
    out[(int64_t)p * o0 + (int64_t)ju * o1 + (int64_t)jv * o2] = val;
}

// shapes: small [V][nu][nv], large [V][Eu][Ev] with even, non-negative padding
int32_t check_pair(const ap_array3* sd, const void* sp, const ap_array3* ld, const void* lp, const void* in,
                   const void* out) {
    int32_t st = ap_check_array3(sd, sp, "projections", false, false, false);
    if (st) return st;
    if ((st = ap_check_array3(ld, lp, "expanded buffer", false, false, false))) return st;
    AP_REQUIRE(in != out || array3_numel(sd) + array3_numel(ld) == 0, "in and out must not overlap");
    AP_REQUIRE(sd->shape[0] == ld->shape[0], "view counts differ");
    for (int a = 1; a < 3; a++) {
        AP_REQUIRE(ld->shape[a] >= sd->shape[a], "the expanded buffer is smaller than the projections");
        AP_REQUIRE((ld->shape[a] - sd->shape[a]) % 2 == 0, "odd padding along axis %d", a);
    }
    AP_REQUIRE(sd->shape[0] <= 65535u, "at most 65535 views per call");
    return AP_OK;
}

}  // namespace

extern "C" int32_t ap_fbp_rotate_weight(const ap_array3* in_desc, const void* in, const ap_array3* out_desc,
                                        void* out, const float* angles, double pix_over_L, double zeta_max,
                                        int32_t mode, void* stream) {
    int32_t st = check_pair(in_desc, in, out_desc, out, in, out);
    if (st) return st;
    AP_REQUIRE(mode == AP_FBP_GBC || mode == AP_FBP_FAN || mode == AP_FBP_FDK, "unknown FBP mode %d", mode);
    AP_REQUIRE(pix_over_L > 0.0, "pix_over_L must be positive");
    const uint32_t V = in_desc->shape[0], Eu = out_desc->shape[1], Ev = out_desc->shape[2];
    if (V == 0 || Eu == 0 || Ev == 0) return AP_OK;
    dim3 grid((Ev + kLanesV - 1) / kLanesV, (Eu + kRowsU - 1) / kRowsU, V);
    hipLaunchKernelGGL(k_fbp_rotate_weight, grid, dim3(kLanesV, kRowsU, 1), 0, (hipStream_t)stream,
                       in_desc->shape[1], in_desc->shape[2], Eu, Ev, (const float*)in, in_desc->strides[0],
                       in_desc->strides[1], in_desc->strides[2], (float*)out, out_desc->strides[0],
                       out_desc->strides[1], out_desc->strides[2], angles, pix_over_L, zeta_max, (int)mode);
    return ap_check_launch("fbp_rotate_weight");
}

extern "C" int32_t ap_fbp_unrotate(const ap_array3* in_desc, const void* in, const ap_array3* out_desc, void* out,
                                   const float* angles, void* stream) {
    int32_t st = check_pair(out_desc, out, in_desc, in, in, out);
    if (st) return st;
    const uint32_t V = out_desc->shape[0], nu = out_desc->shape[1], nv = out_desc->shape[2];
    if (V == 0 || nu == 0 || nv == 0) return AP_OK;
    dim3 grid((nv + kLanesV - 1) / kLanesV, (nu + kRowsU - 1) / kRowsU, V);
    hipLaunchKernelGGL(k_fbp_unrotate, grid, dim3(kLanesV, kRowsU, 1), 0, (hipStream_t)stream, nu, nv,
                       in_desc->shape[1], in_desc->shape[2], (const float*)in, in_desc->strides[0],
                       in_desc->strides[1], in_desc->strides[2], (float*)out, out_desc->strides[0],
                       out_desc->strides[1], out_desc->strides[2], angles);
    return ap_check_launch("fbp_unrotate");
}
