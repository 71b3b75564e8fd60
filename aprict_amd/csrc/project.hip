// project.hip -- ray-driven Siddon forward projector (ap_project_voxel) and
// its matched adjoint (ap_backproject_matched).
//
// Thread mapping: one thread per detector pixel; a wavefront holds 64
// consecutive pixels along v (the fastest sinogram axis, and z, the fastest
// volume axis, for the circular geometries this path serves), so the
// per-step volume gathers of a wave fall in a few cache lines and the
// sinogram stores are fully coalesced.
#include "siddon.h"

namespace {

constexpr int kWaveV = 64;   // pixels along v per wave
constexpr int kRowsU = 4;    // waves (rows along u) per block

struct SinoCfg {
    uint32_t V, nu, nv;
    int64_t s0, s1, s2;      // element strides
};

template <typename Tv, typename Tp, typename IdxT>
__global__ __launch_bounds__(kWaveV* kRowsU) void k_project_voxel(
    RayCfg rc, SinoCfg sc, float mba, uint32_t ssu, uint32_t ssv,
    const ap_proj_spec* __restrict__ specs, const Tv* __restrict__ vol,
    IdxT vsx, IdxT vsy, IdxT vsz, Tp* __restrict__ sino, const Tp* __restrict__ b) {
    const uint32_t k = blockIdx.x * kWaveV + threadIdx.x;
    const uint32_t j = blockIdx.y * kRowsU + threadIdx.y;
    const uint32_t p = blockIdx.z;
    if (k >= sc.nv || j >= sc.nu) return;
    const ViewGeom g = load_view(specs, p);
    const d3 S = to_vox(rc, g.S);
    float tot = 0.f;
    for (uint32_t su = 0; su < ssu; su++) {
        float acc = 0.f;
        for (uint32_t sv = 0; sv < ssv; sv++) {
            const double ju = (double)j - 0.5 + ((double)su + 0.5) / (double)ssu;
            const double kv = (double)k - 0.5 + ((double)sv + 0.5) / (double)ssv;
            const d3 Q = to_vox(rc, pixel_point(g, ju, kv));
            Ray R;
            if (ray_setup(rc, S, Q, R))
                acc += ray_integral<Tv, IdxT>(vol, R, vsx, vsy, vsz, rc.maxit);
        }
        tot += ssv > 1 ? acc / (float)ssv : acc;
    }
    if (ssu > 1) tot /= (float)ssu;
    tot *= (float)(1.0 / rc.inv_vw);   // voxel-length units -> physical
    const int64_t o = (int64_t)p * sc.s0 + (int64_t)j * sc.s1 + (int64_t)k * sc.s2;
    const float base = to_f32(b ? b[o] : sino[o]);
    sino[o] = from_f32<Tp>(__builtin_fmaf(mba, tot, base));
}

// Adjoint by scatter: every sample adds weight*y into the volume with a
// hardware float atomic.  vol must be F32.
template <typename Tp, typename Tm, typename IdxT>
__global__ __launch_bounds__(kWaveV* kRowsU) void k_backproject_matched_atomic(
    RayCfg rc, SinoCfg sc, float scale, uint32_t ssu, uint32_t ssv,
    const ap_proj_spec* __restrict__ specs, float* __restrict__ vol,
    IdxT vsx, IdxT vsy, IdxT vsz, const Tm* __restrict__ mult,
    IdxT msx, IdxT msy, IdxT msz, const Tp* __restrict__ sino) {
    const uint32_t k = blockIdx.x * kWaveV + threadIdx.x;
    const uint32_t j = blockIdx.y * kRowsU + threadIdx.y;
    const uint32_t p = blockIdx.z;
    if (k >= sc.nv || j >= sc.nu) return;
    const int64_t o = (int64_t)p * sc.s0 + (int64_t)j * sc.s1 + (int64_t)k * sc.s2;
    const float y = to_f32(sino[o]);
    if (y == 0.f) return;
    const float wy = y * scale;
    const ViewGeom g = load_view(specs, p);
    const d3 S = to_vox(rc, g.S);
    for (uint32_t su = 0; su < ssu; su++)
        for (uint32_t sv = 0; sv < ssv; sv++) {
            const double ju = (double)j - 0.5 + ((double)su + 0.5) / (double)ssu;
            const double kv = (double)k - 0.5 + ((double)sv + 0.5) / (double)ssv;
            const d3 Q = to_vox(rc, pixel_point(g, ju, kv));
            Ray R;
            if (!ray_setup(rc, S, Q, R)) continue;
            // walk the ray in both the volume and the multiplier index spaces
            IdxT offm = (IdxT)R.idx[0] * msx + (IdxT)R.idx[1] * msy + (IdxT)R.idx[2] * msz;
            const IdxT dmx = R.sg[0] >= 0 ? msx : -msx;
            const IdxT dmy = R.sg[1] >= 0 ? msy : -msy;
            const IdxT dmz = R.sg[2] >= 0 ? msz : -msz;
            IdxT offv = (IdxT)R.idx[0] * vsx + (IdxT)R.idx[1] * vsy + (IdxT)R.idx[2] * vsz;
            const IdxT dvx = R.sg[0] >= 0 ? vsx : -vsx;
            const IdxT dvy = R.sg[1] >= 0 ? vsy : -vsy;
            const IdxT dvz = R.sg[2] >= 0 ? vsz : -vsz;
            const float gx = (float)R.sg[0], gy = (float)R.sg[1], gz = (float)R.sg[2];
            float tx = R.tn[0], ty = R.tn[1], tz = R.tn[2];
            float px = R.pn[0], py = R.pn[1], pz = R.pn[2];
            float tprev = R.t_enter;
            for (int it = 0; it < rc.maxit; ++it) {
                const float tmn = fminf(fminf(tx, ty), tz);
                const float w = fminf(tmn, R.t_exit) - tprev;
                const float m = mult ? to_f32(mult[offm]) : 1.f;
                atomicAdd(vol + offv, w * wy * m);
                if (tmn >= R.t_exit) break;
                tprev = tmn;
                if (tx == tmn) {
                    px += gx; tx = tcross(px, R.r[0], R.c[0]); offv += dvx; offm += dmx;
                } else if (ty == tmn) {
                    py += gy; ty = tcross(py, R.r[1], R.c[1]); offv += dvy; offm += dmy;
                } else {
                    pz += gz; tz = tcross(pz, R.r[2], R.c[2]); offv += dvz; offm += dmz;
                }
            }
        }
}

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
    sc.s0 = sd->strides[0];
    sc.s1 = sd->strides[1];
    sc.s2 = sd->strides[2];
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
                                    const ap_array3* sino_desc, void* sino, const void* b,
                                    void* stream) {
    AP_REQUIRE(specs, "specs is NULL");
    int32_t st = ap_check_array3(vol_desc, vol, "volume", true, true, false);
    if (st) return st;
    SinoCfg sc;
    if ((st = check_sino(sino_desc, sino, num_projs, sc, false))) return st;
    RayCfg rc;
    if ((st = make_ray_cfg(cfg, chunk, vol_desc, rc))) return st;
    if (num_projs == 0 || sc.nu == 0 || sc.nv == 0) return AP_OK;
    AP_REQUIRE(num_projs <= 65535u, "at most 65535 views per call");
    const int64_t maxoff = array3_max_offset(vol_desc);
    const bool idx64 = maxoff >= (int64_t)INT32_MAX;
    dim3 block(kWaveV, kRowsU, 1);
    dim3 grid((sc.nv + kWaveV - 1) / kWaveV, (sc.nu + kRowsU - 1) / kRowsU, num_projs);
    hipStream_t s = (hipStream_t)stream;
    const float mba = cfg->multiply_before_adding;
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
                    if (idx64) {
                        hipLaunchKernelGGL((k_project_voxel<Tv, Tp, int64_t>), grid, block, 0, s, rc, sc,
                                           mba, cfg->ssu, cfg->ssv, specs, (const Tv*)vol,
                                           vol_desc->strides[0], vol_desc->strides[1],
                                           vol_desc->strides[2], (Tp*)sino, (const Tp*)b);
                    } else {
                        hipLaunchKernelGGL((k_project_voxel<Tv, Tp, int32_t>), grid, block, 0, s, rc, sc,
                                           mba, cfg->ssu, cfg->ssv, specs, (const Tv*)vol,
                                           (int32_t)vol_desc->strides[0], (int32_t)vol_desc->strides[1],
                                           (int32_t)vol_desc->strides[2], (Tp*)sino, (const Tp*)b);
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
                                          float alpha, void* stream) {
    AP_REQUIRE(specs, "specs is NULL");
    int32_t st = ap_check_array3(vol_desc, vol, "volume", false, false, false);
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
    AP_REQUIRE(num_projs <= 65535u, "at most 65535 views per call");
    const int64_t maxoff = array3_max_offset(vol_desc);
    const int64_t maxoffm = multiplier ? array3_max_offset(mult_desc) : 0;
    const bool idx64 = maxoff >= (int64_t)INT32_MAX || maxoffm >= (int64_t)INT32_MAX;
    const float scale = (float)((double)alpha * cfg->multiply_before_adding /
                                ((double)cfg->ssu * cfg->ssv) * chunk->voxwid);
    dim3 block(kWaveV, kRowsU, 1);
    dim3 grid((sc.nv + kWaveV - 1) / kWaveV, (sc.nu + kRowsU - 1) / kRowsU, num_projs);
    hipStream_t s = (hipStream_t)stream;
    const int32_t mdt = multiplier ? mult_desc->dtype : AP_F32;
    return dispatch_dtype(sino_desc->dtype, [&](auto tp) -> int32_t {
        using Tp = decltype(tp);
        if constexpr (std::is_same<Tp, double>::value || std::is_same<Tp, bf16_t>::value) {
            return ap_fail(AP_ERR_UNSUPPORTED, "sinogram dtype must be F32 or F16");
        } else {
            return dispatch_dtype(mdt, [&](auto tm) -> int32_t {
                using Tm = decltype(tm);
                if constexpr (std::is_same<Tm, double>::value) {
                    return ap_fail(AP_ERR_UNSUPPORTED, "float64 multiplier");
                } else {
                    const int64_t* ms = multiplier ? mult_desc->strides : vol_desc->strides;
                    if (idx64) {
                        hipLaunchKernelGGL((k_backproject_matched_atomic<Tp, Tm, int64_t>), grid, block, 0, s,
                                           rc, sc, scale, cfg->ssu, cfg->ssv, specs, (float*)vol,
                                           vol_desc->strides[0], vol_desc->strides[1],
                                           vol_desc->strides[2], (const Tm*)multiplier, ms[0], ms[1],
                                           ms[2], (const Tp*)sino);
                    } else {
                        hipLaunchKernelGGL((k_backproject_matched_atomic<Tp, Tm, int32_t>), grid, block, 0, s,
                                           rc, sc, scale, cfg->ssu, cfg->ssv, specs, (float*)vol,
                                           (int32_t)vol_desc->strides[0], (int32_t)vol_desc->strides[1],
                                           (int32_t)vol_desc->strides[2], (const Tm*)multiplier,
                                           (int32_t)ms[0], (int32_t)ms[1], (int32_t)ms[2],
                                           (const Tp*)sino);
                    }
                    return ap_check_launch("backproject_matched");
                }
            });
        }
    });
}
