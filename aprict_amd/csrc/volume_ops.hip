// volume_ops.hip -- element-wise operations on strided volumes / sinograms
// (the texture kernels of cuda/cusrc_templates/texture_operations.cu:124-369),
// the FP64-accumulated inner product, the isotropic-TV gradient and the fused
// SIRT iteration.
#include "ap_common.h"

namespace {

struct Shape3 {
    uint32_t n0, n1, n2;
    int64_t s0, s1, s2;
};

Shape3 shape_of(const ap_array3* d) {
    return Shape3{d->shape[0], d->shape[1], d->shape[2], d->strides[0], d->strides[1], d->strides[2]};
}

__device__ __forceinline__ int64_t offset_of(const Shape3& s, uint64_t n, uint32_t& i, uint32_t& j,
                                             uint32_t& k) {
    k = (uint32_t)(n % s.n2);
    const uint64_t r = n / s.n2;
    j = (uint32_t)(r % s.n1);
    i = (uint32_t)(r / s.n1);
    return (int64_t)i * s.s0 + (int64_t)j * s.s1 + (int64_t)k * s.s2;
}

constexpr int kThreads = 256;

uint32_t grid_for(uint64_t n) {
    uint64_t g = (n + kThreads - 1) / kThreads;
    return (uint32_t)(g > 65536 * 16 ? 65536 * 16 : (g == 0 ? 1 : g));
}

template <typename T>
__global__ void k_fill(Shape3 s, T* a, float v, uint64_t n) {
    for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < n;
         q += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t i, j, k;
        a[offset_of(s, q, i, j, k)] = from_f32<T>(v);
    }
}

template <typename T>
__global__ void k_fill64(Shape3 s, T* a, double v, uint64_t n) {
    for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < n;
         q += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t i, j, k;
        a[offset_of(s, q, i, j, k)] = (T)v;
    }
}

template <typename Ta, typename Tb>
__global__ void k_axpby(Shape3 sa, Ta* a, double alpha, Shape3 sb, const Tb* b, double beta,
                        uint64_t n) {
    for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < n;
         q += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t i, j, k;
        const int64_t oa = offset_of(sa, q, i, j, k);
        const int64_t ob = (int64_t)i * sb.s0 + (int64_t)j * sb.s1 + (int64_t)k * sb.s2;
        const double va = to_f64(a[oa]);
        const double vb = b ? to_f64(b[ob]) : 0.0;
        const double r = alpha * va + beta * vb;
        if constexpr (std::is_same<Ta, double>::value) a[oa] = r;
        else a[oa] = from_f32<Ta>((float)r);
    }
}

template <typename Ta, typename Tb>
__global__ void k_mul(Shape3 sa, Ta* a, Shape3 sb, const Tb* b, uint64_t n) {
    for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < n;
         q += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t i, j, k;
        const int64_t oa = offset_of(sa, q, i, j, k);
        const int64_t ob = (int64_t)i * sb.s0 + (int64_t)j * sb.s1 + (int64_t)k * sb.s2;
        if constexpr (std::is_same<Ta, double>::value) a[oa] = a[oa] * to_f64(b[ob]);
        else a[oa] = from_f32<Ta>(to_f32(a[oa]) * to_f32(b[ob]));
    }
}

// texture_operations.cu:164-185 (1/(v + eps exp(-v/eps)), zeros optional)
template <typename T>
__global__ void k_invert(Shape3 s, T* a, int ignore_zeros, float eps, uint64_t n) {
    for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < n;
         q += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t i, j, k;
        const int64_t o = offset_of(s, q, i, j, k);
        float v = to_f32(a[o]);
        if (ignore_zeros && v == 0.f) continue;
        if (eps > 0.f) v = v + eps * expf(-v / eps);
        a[o] = from_f32<T>(1.f / v);
    }
}

// texture_operations.cu:137-162: raised-cosine taper outside relrad_start of
// the x/y-inscribed cylinder (relative radius 1 at the inscribed circle)
template <typename T>
__global__ void k_softmask(Shape3 s, T* a, float r0, uint64_t n) {
    const float cx = 0.5f * (float)(s.n0 - 1), cy = 0.5f * (float)(s.n1 - 1);
    const float inv = 2.f / (float)min(s.n0, s.n1);
    for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < n;
         q += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t i, j, k;
        const int64_t o = offset_of(s, q, i, j, k);
        const float dx = (float)i - cx, dy = (float)j - cy;
        float rr = sqrtf(dx * dx + dy * dy) * inv;
        if (rr < r0) continue;
        float m;
        if (r0 >= 1.f) {
            m = 0.f;
        } else {
            rr = fminf((rr - r0) / (1.f - r0), 1.f);
            m = fminf(1.f, fmaxf(0.f, 0.5f * (1.f + cosf(rr * 3.14159265358979f))));
        }
        a[o] = from_f32<T>(to_f32(a[o]) * m);
    }
}

__global__ void k_zero_f64(double* out) { *out = 0.0; }

template <typename Ta, typename Tb>
__global__ __launch_bounds__(kThreads) void k_dot(Shape3 sa, const Ta* a, Shape3 sb, const Tb* b,
                                                  uint64_t n, double* out) {
    double acc = 0.0;
    for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < n;
         q += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t i, j, k;
        const int64_t oa = offset_of(sa, q, i, j, k);
        const int64_t ob = (int64_t)i * sb.s0 + (int64_t)j * sb.s1 + (int64_t)k * sb.s2;
        acc += to_f64(a[oa]) * to_f64(b[ob]);
    }
    // wave reduction, then across the block's waves
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    __shared__ double part[kThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kThreads / 64; w++) t += part[w];
        atomicAdd(out, t);
    }
}

// d/dx_i of sum_j sqrt(|D x_j|^2 + eps^2), D = forward differences with
// zero flux at the far faces; x stored F32, F16 or BF16, arithmetic FP32
template <typename Tx>
__global__ __launch_bounds__(kThreads) void k_tv_grad(Shape3 sx, const Tx* x, Shape3 sg,
                                                      float* g, float weight, float eps, uint64_t n) {
    for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < n;
         q += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t i, j, k;
        offset_of(sx, q, i, j, k);
        auto X = [&](int64_t ii, int64_t jj, int64_t kk) -> float {
            return to_f32(x[ii * sx.s0 + jj * sx.s1 + kk * sx.s2]);
        };
        auto grad_at = [&](int64_t ii, int64_t jj, int64_t kk, float& a, float& b, float& c) {
            const float v = X(ii, jj, kk);
            a = ii + 1 < (int64_t)sx.n0 ? X(ii + 1, jj, kk) - v : 0.f;
            b = jj + 1 < (int64_t)sx.n1 ? X(ii, jj + 1, kk) - v : 0.f;
            c = kk + 1 < (int64_t)sx.n2 ? X(ii, jj, kk + 1) - v : 0.f;
        };
        float a, b, c;
        grad_at(i, j, k, a, b, c);
        float r = -(a + b + c) / sqrtf(a * a + b * b + c * c + eps * eps);
        if (i > 0) {
            grad_at(i - 1, j, k, a, b, c);
            r += a / sqrtf(a * a + b * b + c * c + eps * eps);
        }
        if (j > 0) {
            grad_at(i, j - 1, k, a, b, c);
            r += b / sqrtf(a * a + b * b + c * c + eps * eps);
        }
        if (k > 0) {
            grad_at(i, j, k - 1, a, b, c);
            r += c / sqrtf(a * a + b * b + c * c + eps * eps);
        }
        const int64_t og = (int64_t)i * sg.s0 + (int64_t)j * sg.s1 + (int64_t)k * sg.s2;
        g[og] += weight * r;
    }
}

// sum_j sqrt(|D x_j|^2 + eps^2), accumulated in FP64 into *out
template <typename Tx>
__global__ __launch_bounds__(kThreads) void k_tv_value(Shape3 sx, const Tx* x, float eps, uint64_t n,
                                                       double* out) {
    double acc = 0.0;
    for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < n;
         q += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t i, j, k;
        const int64_t o = offset_of(sx, q, i, j, k);
        const float v = to_f32(x[o]);
        const float a = i + 1 < sx.n0 ? to_f32(x[o + sx.s0]) - v : 0.f;
        const float b = j + 1 < sx.n1 ? to_f32(x[o + sx.s1]) - v : 0.f;
        const float c = k + 1 < sx.n2 ? to_f32(x[o + sx.s2]) - v : 0.f;
        acc += (double)sqrtf(a * a + b * b + c * c + eps * eps);
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    __shared__ double part[kThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kThreads / 64; w++) t += part[w];
        atomicAdd(out, t);
    }
}

int32_t same_shape(const ap_array3* a, const ap_array3* b) {
    for (int q = 0; q < 3; q++)
        if (a->shape[q] != b->shape[q]) return ap_fail(AP_ERR_INVALID, "array shapes differ");
    return AP_OK;
}

}  // namespace

extern "C" {

int32_t ap_array_fill(const ap_array3* desc, void* a, double value, void* stream) {
    int32_t st = ap_check_array3(desc, a, "a", true, true, true);
    if (st) return st;
    const uint64_t n = (uint64_t)array3_numel(desc);
    if (n == 0) return AP_OK;
    const Shape3 s = shape_of(desc);
    hipStream_t hs = (hipStream_t)stream;
    return dispatch_dtype(desc->dtype, [&](auto t) -> int32_t {
        using T = decltype(t);
        if constexpr (std::is_same<T, double>::value)
            hipLaunchKernelGGL(k_fill64<double>, dim3(grid_for(n)), dim3(kThreads), 0, hs, s, (double*)a, value, n);
        else
            hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(n)), dim3(kThreads), 0, hs, s, (T*)a, (float)value, n);
        return ap_check_launch("array_fill");
    });
}

int32_t ap_array_axpby(const ap_array3* a_desc, void* a, double alpha, const ap_array3* b_desc,
                       const void* b, double beta, void* stream) {
    int32_t st = ap_check_array3(a_desc, a, "a", true, true, true);
    if (st) return st;
    ap_array3 bd = *a_desc;
    if (b) {
        if ((st = ap_check_array3(b_desc, b, "b", true, true, true))) return st;
        if ((st = same_shape(a_desc, b_desc))) return st;
        bd = *b_desc;
    }
    const uint64_t n = (uint64_t)array3_numel(a_desc);
    if (n == 0) return AP_OK;
    const Shape3 sa = shape_of(a_desc), sb = shape_of(&bd);
    hipStream_t hs = (hipStream_t)stream;
    return dispatch_dtype(a_desc->dtype, [&](auto ta) -> int32_t {
        using Ta = decltype(ta);
        return dispatch_dtype(bd.dtype, [&](auto tb) -> int32_t {
            using Tb = decltype(tb);
            hipLaunchKernelGGL((k_axpby<Ta, Tb>), dim3(grid_for(n)), dim3(kThreads), 0, hs, sa, (Ta*)a,
                               alpha, sb, (const Tb*)b, beta, n);
            return ap_check_launch("array_axpby");
        });
    });
}

int32_t ap_array_scale(const ap_array3* desc, void* a, double factor, void* stream) {
    return ap_array_axpby(desc, a, factor, nullptr, nullptr, 0.0, stream);
}

int32_t ap_array_mul(const ap_array3* a_desc, void* a, const ap_array3* b_desc, const void* b,
                     void* stream) {
    int32_t st = ap_check_array3(a_desc, a, "a", true, true, true);
    if (st) return st;
    if ((st = ap_check_array3(b_desc, b, "b", true, true, true))) return st;
    if ((st = same_shape(a_desc, b_desc))) return st;
    const uint64_t n = (uint64_t)array3_numel(a_desc);
    if (n == 0) return AP_OK;
    const Shape3 sa = shape_of(a_desc), sb = shape_of(b_desc);
    hipStream_t hs = (hipStream_t)stream;
    return dispatch_dtype(a_desc->dtype, [&](auto ta) -> int32_t {
        using Ta = decltype(ta);
        return dispatch_dtype(b_desc->dtype, [&](auto tb) -> int32_t {
            using Tb = decltype(tb);
            hipLaunchKernelGGL((k_mul<Ta, Tb>), dim3(grid_for(n)), dim3(kThreads), 0, hs, sa, (Ta*)a, sb,
                               (const Tb*)b, n);
            return ap_check_launch("array_mul");
        });
    });
}

int32_t ap_array_invert(const ap_array3* desc, void* a, int32_t ignore_zeros, float eps, void* stream) {
    int32_t st = ap_check_array3(desc, a, "a", true, true, false);
    if (st) return st;
    const uint64_t n = (uint64_t)array3_numel(desc);
    if (n == 0) return AP_OK;
    const Shape3 s = shape_of(desc);
    hipStream_t hs = (hipStream_t)stream;
    return dispatch_dtype(desc->dtype, [&](auto t) -> int32_t {
        using T = decltype(t);
        if constexpr (std::is_same<T, double>::value) {
            return ap_fail(AP_ERR_UNSUPPORTED, "float64");
        } else {
            hipLaunchKernelGGL(k_invert<T>, dim3(grid_for(n)), dim3(kThreads), 0, hs, s, (T*)a,
                               (int)ignore_zeros, eps, n);
            return ap_check_launch("array_invert");
        }
    });
}

int32_t ap_volume_softmask_cylinder(const ap_array3* desc, void* a, float relrad_start, void* stream) {
    int32_t st = ap_check_array3(desc, a, "volume", true, true, false);
    if (st) return st;
    AP_REQUIRE(relrad_start >= 0.f && relrad_start <= 1.f, "relrad_start must be in [0,1]");
    const uint64_t n = (uint64_t)array3_numel(desc);
    if (n == 0) return AP_OK;
    const Shape3 s = shape_of(desc);
    hipStream_t hs = (hipStream_t)stream;
    return dispatch_dtype(desc->dtype, [&](auto t) -> int32_t {
        using T = decltype(t);
        if constexpr (std::is_same<T, double>::value) {
            return ap_fail(AP_ERR_UNSUPPORTED, "float64");
        } else {
            hipLaunchKernelGGL(k_softmask<T>, dim3(grid_for(n)), dim3(kThreads), 0, hs, s, (T*)a,
                               relrad_start, n);
            return ap_check_launch("softmask_cylinder");
        }
    });
}

int32_t ap_array_dot(const ap_array3* a_desc, const void* a, const ap_array3* b_desc, const void* b,
                     double* out, void* stream) {
    int32_t st = ap_check_array3(a_desc, a, "a", true, true, true);
    if (st) return st;
    if ((st = ap_check_array3(b_desc, b, "b", true, true, true))) return st;
    if ((st = same_shape(a_desc, b_desc))) return st;
    AP_REQUIRE(out, "out is NULL");
    const uint64_t n = (uint64_t)array3_numel(a_desc);
    hipStream_t hs = (hipStream_t)stream;
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(1), 0, hs, out);
    if (n == 0) return ap_check_launch("array_dot");
    const Shape3 sa = shape_of(a_desc), sb = shape_of(b_desc);
    uint64_t g = (n + kThreads * 8 - 1) / (kThreads * 8);
    if (g > 8192) g = 8192;
    return dispatch_dtype(a_desc->dtype, [&](auto ta) -> int32_t {
        using Ta = decltype(ta);
        return dispatch_dtype(b_desc->dtype, [&](auto tb) -> int32_t {
            using Tb = decltype(tb);
            hipLaunchKernelGGL((k_dot<Ta, Tb>), dim3((uint32_t)g), dim3(kThreads), 0, hs, sa, (const Ta*)a,
                               sb, (const Tb*)b, n, out);
            return ap_check_launch("array_dot");
        });
    });
}

int32_t ap_tv_gradient(const ap_array3* x_desc, const void* x, const ap_array3* g_desc, void* grad,
                       float weight, float eps, void* stream) {
    int32_t st = ap_check_array3(x_desc, x, "x", true, true, false);
    if (st) return st;
    if ((st = ap_check_array3(g_desc, grad, "grad", false, false, false))) return st;
    if ((st = same_shape(x_desc, g_desc))) return st;
    AP_REQUIRE(eps > 0.f, "eps must be positive");
    const uint64_t n = (uint64_t)array3_numel(x_desc);
    if (n == 0) return AP_OK;
    return dispatch_dtype(x_desc->dtype, [&](auto tx) -> int32_t {
        using Tx = decltype(tx);
        if constexpr (std::is_same<Tx, double>::value) {
            return ap_fail(AP_ERR_UNSUPPORTED, "x: float64 not supported here");
        } else {
            hipLaunchKernelGGL((k_tv_grad<Tx>), dim3(grid_for(n)), dim3(kThreads), 0, (hipStream_t)stream,
                               shape_of(x_desc), (const Tx*)x, shape_of(g_desc), (float*)grad, weight, eps, n);
            return ap_check_launch("tv_gradient");
        }
    });
}

int32_t ap_tv_value(const ap_array3* x_desc, const void* x, float eps, double* out, void* stream) {
    int32_t st = ap_check_array3(x_desc, x, "x", true, true, false);
    if (st) return st;
    AP_REQUIRE(eps > 0.f, "eps must be positive");
    AP_REQUIRE(out, "out is NULL");
    const uint64_t n = (uint64_t)array3_numel(x_desc);
    hipStream_t hs = (hipStream_t)stream;
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(1), 0, hs, out);
    if (n == 0) return ap_check_launch("tv_value");
    uint64_t gr = (n + kThreads * 8 - 1) / (kThreads * 8);
    if (gr > 8192) gr = 8192;
    return dispatch_dtype(x_desc->dtype, [&](auto tx) -> int32_t {
        using Tx = decltype(tx);
        if constexpr (std::is_same<Tx, double>::value) {
            return ap_fail(AP_ERR_UNSUPPORTED, "x: float64 not supported here");
        } else {
            hipLaunchKernelGGL((k_tv_value<Tx>), dim3((uint32_t)gr), dim3(kThreads), 0, hs, shape_of(x_desc),
                               (const Tx*)x, eps, n, out);
            return ap_check_launch("tv_value");
        }
    });
}

int32_t ap_sirt_iterate(const ap_config_cb_backproject* bp_cfg, const ap_proj_spec* specs,
                        uint32_t num_projs, const ap_vol_tex_chunk_spec* chunk,
                        const ap_array3* vol_desc, void* x, const ap_array3* m_desc, const void* M,
                        const ap_array3* b_desc, const void* b, const ap_array3* r_desc, void* r,
                        void* stream) {
    AP_REQUIRE(bp_cfg && b && r && b_desc && r_desc, "NULL config, measurements or residual buffer");
    AP_REQUIRE(b != r, "residual buffer must not alias the measurements");
    for (int a = 0; a < 3; a++)
        AP_REQUIRE(r_desc->shape[a] == b_desc->shape[a], "residual shape != measurement shape along axis %d", a);
    AP_REQUIRE(r_desc->dtype == b_desc->dtype, "residual dtype != measurement dtype");
    // r = b - A x, written once by the projector epilogue (b and r each with its own strides)
    const ap_config_cb_project fp = {1.f, -1.f, 1u, 1u};
    int32_t st = ap_project_voxel(&fp, specs, num_projs, chunk, vol_desc, x, r_desc, r, b_desc, b, stream);
    if (st) return st;
    // x += M (.) BP(r)
    return ap_backproject(bp_cfg, specs, num_projs, chunk, vol_desc, x, m_desc, M, r_desc, r, nullptr, nullptr,
                          stream);
}

}  // extern "C"
