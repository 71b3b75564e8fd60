// ap_common.h -- shared device/host helpers of libaprict_hip (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/aprict_hip.h"

// ---------------------------------------------------------------------------
// error reporting (ap_last_error is thread-local)
// ---------------------------------------------------------------------------
int32_t ap_fail(int32_t code, const char* fmt, ...);
// selected launch-shape variant of a tunable operation (AP_TUNE_*)
int32_t ap_tune_variant(int32_t op);
int32_t ap_check_launch(const char* what);

#define AP_REQUIRE(cond, ...)                                   \
    do {                                                        \
        if (!(cond)) return ap_fail(AP_ERR_INVALID, __VA_ARGS__); \
    } while (0)

// ---------------------------------------------------------------------------
// element access for the storage types (F32 / F16 / BF16 / F64)
// ---------------------------------------------------------------------------
typedef uint16_t bf16_t;   // raw bfloat16 bits
typedef _Float16 f16_t;

__device__ __forceinline__ float to_f32(float x) { return x; }
__device__ __forceinline__ float to_f32(double x) { return (float)x; }
__device__ __forceinline__ float to_f32(f16_t x) { return (float)x; }
__device__ __forceinline__ float to_f32(bf16_t x) {
    return __uint_as_float(((uint32_t)x) << 16);
}

template <typename T> __device__ __forceinline__ T from_f32(float x);
template <> __device__ __forceinline__ float from_f32<float>(float x) { return x; }
template <> __device__ __forceinline__ double from_f32<double>(float x) { return (double)x; }
template <> __device__ __forceinline__ f16_t from_f32<f16_t>(float x) { return (f16_t)x; }
template <> __device__ __forceinline__ bf16_t from_f32<bf16_t>(float x) {
    // round to nearest even; NaN stays NaN
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}

// double-precision load/store used by reductions and the FP64 paths
__device__ __forceinline__ double to_f64(double x) { return x; }
template <typename T> __device__ __forceinline__ double to_f64(T x) { return (double)to_f32(x); }

inline int dtype_size(int32_t dt) {
    switch (dt) {
    case AP_F32: return 4;
    case AP_F16: return 2;
    case AP_BF16: return 2;
    case AP_F64: return 8;
    default: return 0;
    }
}

// Largest element offset reachable through a descriptor (for bounds of
// 32-bit indexing); -1 for a malformed descriptor.
inline int64_t array3_max_offset(const ap_array3* d) {
    int64_t m = 0;
    for (int a = 0; a < 3; a++) {
        if (d->shape[a] == 0) return 0;
        if (d->strides[a] < 0) return -1;
        m += (int64_t)(d->shape[a] - 1) * d->strides[a];
    }
    return m;
}

inline int64_t array3_numel(const ap_array3* d) {
    return (int64_t)d->shape[0] * d->shape[1] * d->shape[2];
}

int32_t ap_check_array3(const ap_array3* d, const void* ptr, const char* name,
                        bool allow_f16, bool allow_bf16, bool allow_f64);

// ---------------------------------------------------------------------------
// dispatch a functor over a storage dtype
// ---------------------------------------------------------------------------
template <typename F>
inline int32_t dispatch_dtype(int32_t dt, F&& f) {
    switch (dt) {
    case AP_F32: return f(float{});
    case AP_F16: return f(f16_t{});
    case AP_BF16: return f(bf16_t{});
    case AP_F64: return f(double{});
    default: return ap_fail(AP_ERR_UNSUPPORTED, "unsupported dtype %d", dt);
    }
}

// ---------------------------------------------------------------------------
// small vector helpers
// ---------------------------------------------------------------------------
struct d3 { double x, y, z; };
__host__ __device__ __forceinline__ d3 mkd3(double x, double y, double z) { return d3{x, y, z}; }
__host__ __device__ __forceinline__ d3 operator+(d3 a, d3 b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ __forceinline__ d3 operator-(d3 a, d3 b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ __forceinline__ d3 operator*(d3 a, double s) { return d3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ __forceinline__ double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ d3 cross(d3 a, d3 b) {
    return d3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ __forceinline__ d3 fv(ap_fvec3 a) { return d3{a.x, a.y, a.z}; }

// Load a proj_spec field by field (the 60-byte stride is not 16-byte aligned).
struct ViewGeom {
    d3 S, D, U, V;   // source, detector centre, pixel vectors
    uint32_t nu, nv;
};

__device__ __forceinline__ ViewGeom load_view(const ap_proj_spec* __restrict__ specs, uint32_t p) {
    const ap_proj_spec* s = specs + p;
    ViewGeom g;
    g.S = fv(s->sourcepos);
    g.D = fv(s->dspec.pos);
    g.U = fv(s->dspec.u);
    g.V = fv(s->dspec.v);
    g.nu = s->dspec.nu;
    g.nv = s->dspec.nv;
    return g;
}

constexpr float AP_BIG = 3.0e38f;
