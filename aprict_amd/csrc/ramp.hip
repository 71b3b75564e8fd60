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
// Block: 64 lanes along v (coalesced) x 4 rows; a block owns 64 outputs j of
// one view.  The input rows m are staged through LDS kChunk at a time (eight
// loads in flight per thread), each chunk with the taps it needs (so nu is not
// limited by LDS); the taps a wave reads for one m are uniform (LDS broadcast).
//  - Ram-Lak (ramp_kernels.py:43-59) is zero at even nonzero offsets, so each
//    thread owns 16 outputs of one parity (j, j+2, ..., j+30) and visits only
//    the rows m of the other parity, where every tap is an odd one; the
//    centre tap (m = j) is added at the end.  Half the multiplies of the
//    full convolution.
//  - the other kernels (sinc ramp, Hilbert taps) have no zero taps: each
//    thread owns 16 consecutive outputs and visits every row.
#include "ap_common.h"

namespace {

constexpr int kLanes = 64, kRows = 4, kPerThread = 16, kChunk = 128;
constexpr int kOutPerBlock = kRows * kPerThread;   // 64 outputs j per block

__device__ __forceinline__ double tap_value(int kernel, long d, double p) {
    const double pi = 3.14159265358979323846;
    if (kernel == AP_RAMP_SINC) return 8.0 / (1.0 - 4.0 * (double)d * (double)d) / p;
    if (kernel == AP_RAMP_HILBERT) return d == 0 ? 0.0 : 2.0 / (double)d;   // independent of p
    if (kernel == AP_RAMP_HILBERT_HALF) return 2.0 / ((double)d + 0.5);
    if (d == 0) return pi * pi / p;
    if (d % 2 == 0) return 0.0;
    return -4.0 / ((double)d * (double)d) / p;
}

template <bool kRamLak>
__global__ __launch_bounds__(kLanes* kRows) void k_ramp_rows(
    uint32_t nu, uint32_t nv, const float* __restrict__ in, int64_t i0, int64_t i1, int64_t i2,
    float* __restrict__ out, int64_t o0, int64_t o1, int64_t o2, double p, int kernel, int fdk) {
    // taps of offsets d = dlo + e, dlo = J0 - (m0 + kChunk - 1), e < kChunk + kOutPerBlock
    __shared__ float taps[kChunk + kOutPerBlock];
    __shared__ float xs[kChunk][kLanes];
    const int tid = threadIdx.y * kLanes + threadIdx.x;
    const uint32_t k = blockIdx.x * kLanes + threadIdx.x;
    const int J0 = (int)(blockIdx.y * kOutPerBlock);
    const uint32_t view = blockIdx.z;
    // this thread's outputs: j_q = jt + q * js
    const int ty = (int)threadIdx.y;
    const int jt = kRamLak ? J0 + 2 * kPerThread * (ty >> 1) + (ty & 1) : J0 + ty * kPerThread;
    const int js = kRamLak ? 2 : 1;
    float acc[kPerThread];
    for (int q = 0; q < kPerThread; q++) acc[q] = 0.f;
    const float vv = (float)(p * ((double)k - 0.5 * (double)(nv - 1)));
    const float* __restrict__ col = in + (int64_t)view * i0 + (int64_t)(k < nv ? k : 0) * i2;
    auto weight = [&](uint32_t m, float x) -> float {
        if (fdk) {
            // the cosine weight with the hardware reciprocal square root (about 1 ulp;
            // the IEEE square root and division are twenty instructions, and every
            // block stages each input of its view again)
            const float uu = (float)(p * ((double)m - 0.5 * (double)(nu - 1)));
            x *= __builtin_amdgcn_rsqf(__builtin_fmaf(uu, uu, __builtin_fmaf(vv, vv, 1.f)));
        }
        return x;
    };
    constexpr int kLd = 8;   // input loads in flight per thread
    for (uint32_t m0 = 0; m0 < nu; m0 += kChunk) {
        const uint32_t nm = min((uint32_t)kChunk, nu - m0);
        const long dlo = (long)J0 - (long)(m0 + kChunk - 1);
        __syncthreads();
        for (int e = tid; e < kChunk + kOutPerBlock; e += kLanes * kRows) taps[e] = (float)tap_value(kernel, dlo + e, p);
        // kLd loads issued before the first is used
        for (uint32_t rb = threadIdx.y; rb < nm; rb += kRows * kLd) {
            // every load unconditional, from a clamped (in-range) row and column,
            // so that none waits for a branch; the writes select
            float v[kLd];
#pragma unroll
            for (int u = 0; u < kLd; u++) {
                const uint32_t r = min(rb + (uint32_t)(u * kRows), nm - 1);
                v[u] = col[(int64_t)(m0 + r) * i1];
            }
#pragma unroll
            for (int u = 0; u < kLd; u++) {
                const uint32_t r = rb + (uint32_t)(u * kRows);
                if (r < nm) xs[r][threadIdx.x] = k < nv ? weight(m0 + r, v[u]) : 0.f;
            }
        }
        __syncthreads();
        // taps[base + q * js] = tap(j_q - m) with base = jt - m - dlo.  Ram-Lak visits
        // the rows of the other parity than jt (every j_q - m odd), the other kernels
        // every row: rows rs apart, rs = js.
        constexpr int rs = kRamLak ? 2 : 1;
        const uint32_t r0 = kRamLak ? (uint32_t)((jt + 1 - (int)m0) & 1) : 0u;
        uint32_t r = r0;
        {
            // From one row to the next the thread's taps slide by one: output q then
            // takes the tap output q - 1 had, and output 0 a new one.  The taps live in
            // a ring of registers, slot (q - s) mod kPerThread for output q at step s,
            // so kPerThread rows at a time need kPerThread new taps (one LDS read per
            // row, not kPerThread).
            int base = jt - (int)(m0 + r) - (int)dlo;
            float t[kPerThread];
#pragma unroll
            for (int q = 0; q < kPerThread; q++) t[q] = taps[base + js * q];
            for (; r + (uint32_t)(rs * (kPerThread - 1)) < nm; r += (uint32_t)(rs * kPerThread)) {
#pragma unroll
                for (int st = 0; st < kPerThread; st++) {
                    const float x = xs[r + rs * st][threadIdx.x];
#pragma unroll
                    for (int q = 0; q < kPerThread; q++)
                        acc[q] = __builtin_fmaf(x, t[(q - st + kPerThread) % kPerThread], acc[q]);
                    // the tap output 0 takes at the next row (an index past the staged
                    // ones only after the last row, where it is not used)
                    t[(kPerThread - 1 - st) % kPerThread] = taps[max(base - js * (st + 1), 0)];
                }
                base -= js * kPerThread;
            }
        }
        for (; r < nm; r += rs) {
            const float x = xs[r][threadIdx.x];
            const int base = jt - (int)(m0 + r) - (int)dlo;
#pragma unroll
            for (int q = 0; q < kPerThread; q++) acc[q] = __builtin_fmaf(x, taps[base + js * q], acc[q]);
        }
    }
    if (k >= nv) return;
    const float centre = kRamLak ? (float)tap_value(AP_RAMP_RAMLAK, 0, p) : 0.f;
    float xc[kPerThread];
#pragma unroll
    for (int q = 0; q < kPerThread; q++) {
        const int j = jt + q * js;
        xc[q] = kRamLak && j < (int)nu ? col[(int64_t)j * i1] : 0.f;
    }
#pragma unroll
    for (int q = 0; q < kPerThread; q++) {
        const int j = jt + q * js;
        if (j >= (int)nu) continue;
        float a = acc[q];
        if (kRamLak) a = __builtin_fmaf(weight((uint32_t)j, xc[q]), centre, a);
        out[(int64_t)view * o0 + (int64_t)j * o1 + (int64_t)k * o2] = a;
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
    AP_REQUIRE(nu < (1u << 30), "nu above 2^30");
    AP_REQUIRE(V <= 65535u, "at most 65535 views per call");
    dim3 grid((nv + kLanes - 1) / kLanes, (nu + kOutPerBlock - 1) / kOutPerBlock, V);
    auto launch = [&](auto kern) {
        hipLaunchKernelGGL(kern, grid, dim3(kLanes, kRows, 1), 0, (hipStream_t)stream, nu, nv, (const float*)in,
                           in_desc->strides[0], in_desc->strides[1], in_desc->strides[2], (float*)out,
                           out_desc->strides[0], out_desc->strides[1], out_desc->strides[2], physpixwid,
                           (int)kernel, (int)fdk_cosine_weight);
    };
    if (kernel == AP_RAMP_RAMLAK) launch(k_ramp_rows<true>);
    else launch(k_ramp_rows<false>);
    return ap_check_launch("ramp_filter_rows");
}
