// ramp.hip -- row-wise ramp filter (ap_ramp_filter_rows): direct linear
// convolution along axis 1 (u) of a [V][nu][nv] sinogram.
//
// The reference filters with an FP32 FFT (cb/ramp_filter.py:139-240, taps of
// ramp_kernels.py:27-94 transformed in FP64); the convolution it implements is
// the "same"-mode linear convolution with those taps, which this kernel
// evaluates directly (no padding, no complex arithmetic, FP32 data with the
// taps rounded once from FP64).  The FDK cosine pre-weight of
// cb/ramp_filter.py:407-436 is fused into the staging of the input.
//
// Block: 64 lanes along v (coalesced) x 4 rows; each thread owns 8
// consecutive outputs j, so the taps a wave reads are uniform (LDS broadcast).
#include "ap_common.h"

namespace {

constexpr int kLanes = 64, kRows = 4, kPerThread = 8, kChunk = 128;
constexpr int kOutPerBlock = kRows * kPerThread;   // 32 outputs j per block
constexpr int kMaxNu = 4096;

__global__ __launch_bounds__(kLanes* kRows) void k_ramp_rows(
    uint32_t nu, uint32_t nv, const float* __restrict__ in, int64_t i0, int64_t i1, int64_t i2,
    float* __restrict__ out, int64_t o0, int64_t o1, int64_t o2, double p, int kernel, int fdk) {
    // tap(n) at taps[n + nu - 1]; valid outputs j < nu only read n = j - m in
    // [-(nu - 1), nu - 1], and the outputs j >= nu of a block's last rows
    // (discarded) read at most kOutPerBlock entries past 2 nu - 1: padding,
    // so the loop needs no range check
    __shared__ float taps[2 * kMaxNu + kOutPerBlock];
    __shared__ float xs[kChunk][kLanes];
    const int tid = threadIdx.y * kLanes + threadIdx.x;
    const uint32_t k = blockIdx.x * kLanes + threadIdx.x;
    const uint32_t jb = blockIdx.y * kOutPerBlock + threadIdx.y * kPerThread;
    const uint32_t view = blockIdx.z;
    const double pi = 3.14159265358979323846;
    for (int n = tid; n < (int)(2 * nu - 1); n += kLanes * kRows) {
        const long d = (long)n - (long)(nu - 1);
        double t;
        if (kernel == AP_RAMP_SINC) t = 8.0 / (1.0 - 4.0 * (double)d * (double)d) / p;
        else if (kernel == AP_RAMP_HILBERT) t = d == 0 ? 0.0 : 2.0 / (double)d;   // independent of p
        else if (kernel == AP_RAMP_HILBERT_HALF) t = 2.0 / ((double)d + 0.5);
        else if (d == 0) t = pi * pi / p;
        else if (d % 2 == 0) t = 0.0;
        else t = -4.0 / ((double)d * (double)d) / p;
        taps[n] = (float)t;
    }
    float acc[kPerThread];
    for (int q = 0; q < kPerThread; q++) acc[q] = 0.f;
    const float vv = (float)(p * ((double)k - 0.5 * (double)(nv - 1)));
    for (uint32_t m0 = 0; m0 < nu; m0 += kChunk) {
        const uint32_t nm = min((uint32_t)kChunk, nu - m0);
        __syncthreads();
        for (uint32_t r = threadIdx.y; r < nm; r += kRows) {
            float x = 0.f;
            if (k < nv) {
                x = in[(int64_t)view * i0 + (int64_t)(m0 + r) * i1 + (int64_t)k * i2];
                if (fdk) {
                    const float uu = (float)(p * ((double)(m0 + r) - 0.5 * (double)(nu - 1)));
                    x *= 1.f / sqrtf(1.f + uu * uu + vv * vv);
                }
            }
            xs[r][threadIdx.x] = x;
        }
        __syncthreads();
        for (uint32_t r = 0; r < nm; r++) {
            const float x = xs[r][threadIdx.x];
            const int base = (int)jb - (int)(m0 + r) + (int)nu - 1;
#pragma unroll
            for (int q = 0; q < kPerThread; q++) acc[q] = __builtin_fmaf(x, taps[base + q], acc[q]);
        }
    }
    if (k >= nv) return;
    for (int q = 0; q < kPerThread; q++) {
        const uint32_t j = jb + q;
        if (j < nu) out[(int64_t)view * o0 + (int64_t)j * o1 + (int64_t)k * o2] = acc[q];
    }
}

}  // namespace

extern "C" int32_t ap_ramp_filter_rows(const ap_array3* in_desc, const void* in,
                                       const ap_array3* out_desc, void* out, double physpixwid,
                                       int32_t kernel, int32_t fdk_cosine_weight, void* stream) {
    int32_t st = ap_check_array3(in_desc, in, "in", false, false, false);
    if (st) return st;
    if ((st = ap_check_array3(out_desc, out, "out", false, false, false))) return st;
    for (int a = 0; a < 3; a++)
        AP_REQUIRE(in_desc->shape[a] == out_desc->shape[a], "in and out shapes differ");
    AP_REQUIRE(in != out || array3_numel(in_desc) == 0, "in and out must not overlap");
    AP_REQUIRE(physpixwid > 0.0, "physpixwid must be positive");
    AP_REQUIRE(kernel == AP_RAMP_RAMLAK || kernel == AP_RAMP_SINC || kernel == AP_RAMP_HILBERT ||
                   kernel == AP_RAMP_HILBERT_HALF,
               "unknown ramp kernel %d", kernel);
    const uint32_t V = in_desc->shape[0], nu = in_desc->shape[1], nv = in_desc->shape[2];
    if (V == 0 || nu == 0 || nv == 0) return AP_OK;
    AP_REQUIRE(nu <= (uint32_t)kMaxNu, "nu above %d", kMaxNu);
    AP_REQUIRE(V <= 65535u, "at most 65535 views per call");
    dim3 grid((nv + kLanes - 1) / kLanes, (nu + kOutPerBlock - 1) / kOutPerBlock, V);
    hipLaunchKernelGGL(k_ramp_rows, grid, dim3(kLanes, kRows, 1), 0, (hipStream_t)stream, nu, nv,
                       (const float*)in, in_desc->strides[0], in_desc->strides[1], in_desc->strides[2],
                       (float*)out, out_desc->strides[0], out_desc->strides[1], out_desc->strides[2],
                       physpixwid, (int)kernel, (int)fdk_cosine_weight);
    return ap_check_launch("ramp_filter_rows");
}
