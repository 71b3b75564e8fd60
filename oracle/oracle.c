/*
 * oracle.c -- CPU restatement of aprict's projection hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only the test suite, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load the libraries built from this file;
 * the product path (aprict_amd) never calls it.
 *
 * Built twice from this one source (oracle/Makefile):
 *   liboracle64.so  REAL = double  -- the FP64 oracle the HIP kernels are
 *                                     checked against;
 *   liboracle32.so  REAL = float   -- the FP32 OpenMP port of the reference
 *                                     semantics, timed as the CPU baseline.
 *
 * Paths cited below are relative to aprict's src/aprict/.  Arrays are dense
 * C-order: volume [Nx][Ny][Nz], sinogram [V][nu][nv].
 *
 * Independence: the forward projector here is the classical Siddon
 * algorithm (merge of the three sorted plane-crossing sequences, voxel found
 * from the segment midpoint), not the DDA the GPU kernel runs, and
 * or_slab_length() is a third, brute-force route to the same weights.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/aprict_hip.h"

#ifndef OR_REAL
#define OR_REAL double
#endif
typedef OR_REAL real;

#ifdef OR_FP32
#define OR_SUFFIX(name) name##_f32
#else
#define OR_SUFFIX(name) name
#endif

typedef struct { real x, y, z; } v3;

static inline v3 mk(real x, real y, real z) { v3 r = {x, y, z}; return r; }
static inline v3 vf(ap_fvec3 a) { return mk((real)a.x, (real)a.y, (real)a.z); }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scl(v3 a, real s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline real dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline real nrm(v3 a) { return (real)sqrt((double)dot(a, a)); }
static inline v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline real comp(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
static inline real rmax(real a, real b) { return a > b ? a : b; }
static inline real rmin(real a, real b) { return a < b ? a : b; }

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* Pixel-centre position of a flat-panel detector
 * (CB_project_texture_chunked.cu:308-324). */
static v3 det_point(const ap_det_spec* d, real ju, real kv) {
    v3 p = vf(d->pos);
    p = add(p, scl(vf(d->u), ju - (real)(d->nu - 1) / 2));
    p = add(p, scl(vf(d->v), kv - (real)(d->nv - 1) / 2));
    return p;
}

/* ======================================================================
 * Forward projection: exact line integral through a voxel chunk.
 * Semantics: CB_project_texture_chunked.cu:111-306 (line integral of the
 * piecewise-constant chunk along the finite segment S->Q) and :327-412
 * (sub-pixel sampling, averaging, multiply_before_adding, accumulation).
 * ====================================================================== */

/* Line integral in voxel units: S, Q already in voxel coordinates (corner of
 * voxel 000 at the origin).  lo/hi: chunk box.  Returns sum(len_i * f_i) with
 * len in voxel units. */
static real siddon_voxunits(v3 S, v3 Q, const int lo[3], const int hi[3],
                            const uint32_t dims[3], const real* vol) {
    v3 d = sub(Q, S);
    real t0 = 0, t1 = 1;
    int fixed_idx[3];
    for (int a = 0; a < 3; a++) {
        real s = comp(S, a), da = comp(d, a);
        if (da == 0) {
            /* parallel to the planes of this axis: inside iff lo < s <= hi,
             * in voxel ceil(s)-1 (the reference's dir==0 convention,
             * CB_project_texture_chunked.cu:189-203,244-246) */
            if (!(s > lo[a] && s <= hi[a])) return 0;
            fixed_idx[a] = (int)ceil((double)s) - 1;
        } else {
            real ta = (lo[a] - s) / da, tb = (hi[a] - s) / da;
            t0 = rmax(t0, rmin(ta, tb));
            t1 = rmin(t1, rmax(ta, tb));
            fixed_idx[a] = -1;
        }
    }
    if (!(t0 < t1)) return 0;
    /* per-axis increasing sequences of interior plane crossings */
    int n[3], next[3], step[3];
    for (int a = 0; a < 3; a++) {
        real da = comp(d, a);
        n[a] = 0;
        if (da == 0) continue;
        step[a] = da > 0 ? 1 : -1;
        /* first interior plane after t0 in the direction of travel */
        real pos0 = comp(S, a) + t0 * da;
        int p = da > 0 ? (int)floor((double)pos0) + 1 : (int)ceil((double)pos0) - 1;
        if (da > 0 && p <= lo[a]) p = lo[a] + 1;
        if (da < 0 && p >= hi[a]) p = hi[a] - 1;
        next[a] = p;
        n[a] = 1;
    }
    real acc = 0, tprev = t0;
    for (;;) {
        /* smallest pending crossing among the three axes */
        real tbest = t1;
        int abest = -1;
        for (int a = 0; a < 3; a++) {
            if (!n[a]) continue;
            int p = next[a];
            if (p <= lo[a] || p >= hi[a]) { n[a] = 0; continue; }
            real t = (p - comp(S, a)) / comp(d, a);
            if (t <= tprev) { next[a] += step[a]; a--; continue; } /* behind: skip */
            if (t < tbest) { tbest = t; abest = a; }
        }
        /* segment [tprev, tbest]: voxel of its midpoint */
        if (tbest > tprev) {
            real tm = (tprev + tbest) / 2;
            int idx[3];
            for (int a = 0; a < 3; a++) {
                if (fixed_idx[a] >= 0) { idx[a] = fixed_idx[a]; continue; }
                int i = (int)floor((double)(comp(S, a) + tm * comp(d, a)));
                if (i < lo[a]) i = lo[a];
                if (i > hi[a] - 1) i = hi[a] - 1;
                idx[a] = i;
            }
            size_t off = ((size_t)idx[0] * dims[1] + idx[1]) * dims[2] + idx[2];
            acc += (tbest - tprev) * vol[off];
            tprev = tbest;
        }
        if (abest < 0) break;
        next[abest] += step[abest];
    }
    return acc * nrm(d);
}

/* Length of the segment S->Q inside the box [lo,hi] (any units): the
 * brute-force slab formula used to build dense system matrices in tests. */
real OR_SUFFIX(or_slab_length)(const double* S3, const double* Q3,
                               const double* lo3, const double* hi3) {
    v3 S = mk(S3[0], S3[1], S3[2]), Q = mk(Q3[0], Q3[1], Q3[2]);
    v3 d = sub(Q, S);
    real t0 = 0, t1 = 1;
    for (int a = 0; a < 3; a++) {
        real s = comp(S, a), da = comp(d, a);
        if (da == 0) {
            if (s < lo3[a] || s > hi3[a]) return 0;
            continue;
        }
        real ta = (lo3[a] - s) / da, tb = (hi3[a] - s) / da;
        t0 = rmax(t0, rmin(ta, tb));
        t1 = rmin(t1, rmax(ta, tb));
    }
    return t1 > t0 ? (t1 - t0) * nrm(d) : 0;
}

typedef struct {
    real svf, voxwid;
    v3 texpos;       /* physical */
    int lo[3], hi[3];
} fwd_geom;

static fwd_geom make_fwd_geom(const ap_config_cb_project* cfg,
                              const ap_vol_tex_chunk_spec* ch) {
    fwd_geom g;
    g.svf = cfg->shrink_volume_factor;
    g.voxwid = ch->voxwid;
    g.texpos = vf(ch->texpos);
    g.lo[0] = (int)ch->offsets.x; g.lo[1] = (int)ch->offsets.y; g.lo[2] = (int)ch->offsets.z;
    g.hi[0] = g.lo[0] + (int)ch->shape.x;
    g.hi[1] = g.lo[1] + (int)ch->shape.y;
    g.hi[2] = g.lo[2] + (int)ch->shape.z;
    return g;
}

static inline v3 to_vox(const fwd_geom* g, v3 p) {
    return scl(sub(scl(p, g->svf), g->texpos), 1 / g->voxwid);
}

/* One pixel of the forward projector (sub-rays averaged, x mba). */
static real fwd_pixel(const ap_config_cb_project* cfg, const fwd_geom* g,
                      const ap_proj_spec* ps, uint32_t j, uint32_t k,
                      const uint32_t dims[3], const real* vol) {
    v3 S = to_vox(g, vf(ps->sourcepos));
    real tot = 0;
    for (uint32_t su = 0; su < cfg->ssu; su++) {
        real acc = 0;
        for (uint32_t sv = 0; sv < cfg->ssv; sv++) {
            real ju = (real)j - (real)0.5 + ((real)su + (real)0.5) / (real)cfg->ssu;
            real kv = (real)k - (real)0.5 + ((real)sv + (real)0.5) / (real)cfg->ssv;
            v3 Q = to_vox(g, det_point(&ps->dspec, ju, kv));
            acc += siddon_voxunits(S, Q, g->lo, g->hi, dims, vol) * g->voxwid;
        }
        tot += acc / (real)cfg->ssv;
    }
    return tot / (real)cfg->ssu * (real)cfg->multiply_before_adding;
}

/* sino[p][j][k] += mba * mean_ss integral   (b == NULL)
 * sino[p][j][k]  = b[p][j][k] + mba * ...    (b != NULL) */
int OR_SUFFIX(or_project_voxel)(const ap_config_cb_project* cfg,
                                const ap_proj_spec* specs, uint32_t nprojs,
                                const ap_vol_tex_chunk_spec* chunk,
                                const uint32_t* dims, const real* vol,
                                real* sino, const real* b, int nthreads) {
    if (!cfg || !specs || !chunk || !dims || !vol || !sino) return AP_ERR_INVALID;
    if (cfg->ssu == 0 || cfg->ssv == 0) return AP_ERR_INVALID;
    set_threads(nthreads);
    fwd_geom g = make_fwd_geom(cfg, chunk);
    const uint32_t nu = specs[0].dspec.nu, nv = specs[0].dspec.nv;
    long long total = (long long)nprojs * nu;
#pragma omp parallel for schedule(dynamic, 4)
    for (long long pj = 0; pj < total; pj++) {
        uint32_t p = (uint32_t)(pj / nu), j = (uint32_t)(pj % nu);
        for (uint32_t k = 0; k < nv; k++) {
            size_t o = ((size_t)p * nu + j) * nv + k;
            real val = fwd_pixel(cfg, &g, &specs[p], j, k, dims, vol);
            sino[o] = (b ? b[o] : sino[o]) + val;
        }
    }
    return AP_OK;
}

/* Transpose of or_project_voxel: vol[i] += alpha*mult[i]*(A^T sino)[i].
 * Built from the same Siddon traversal with the scatter written as a
 * per-ray list; serial over rays (deterministic). */
static void siddon_scatter(v3 S, v3 Q, const int lo[3], const int hi[3],
                           const uint32_t dims[3], real w, real* out) {
    v3 d = sub(Q, S);
    real t0 = 0, t1 = 1;
    int fixed_idx[3];
    for (int a = 0; a < 3; a++) {
        real s = comp(S, a), da = comp(d, a);
        if (da == 0) {
            if (!(s > lo[a] && s <= hi[a])) return;
            fixed_idx[a] = (int)ceil((double)s) - 1;
        } else {
            real ta = (lo[a] - s) / da, tb = (hi[a] - s) / da;
            t0 = rmax(t0, rmin(ta, tb));
            t1 = rmin(t1, rmax(ta, tb));
            fixed_idx[a] = -1;
        }
    }
    if (!(t0 < t1)) return;
    int n[3], next[3], step[3];
    for (int a = 0; a < 3; a++) {
        real da = comp(d, a);
        n[a] = 0;
        if (da == 0) continue;
        step[a] = da > 0 ? 1 : -1;
        real pos0 = comp(S, a) + t0 * da;
        int p = da > 0 ? (int)floor((double)pos0) + 1 : (int)ceil((double)pos0) - 1;
        if (da > 0 && p <= lo[a]) p = lo[a] + 1;
        if (da < 0 && p >= hi[a]) p = hi[a] - 1;
        next[a] = p;
        n[a] = 1;
    }
    real len = nrm(d), tprev = t0;
    for (;;) {
        real tbest = t1;
        int abest = -1;
        for (int a = 0; a < 3; a++) {
            if (!n[a]) continue;
            int p = next[a];
            if (p <= lo[a] || p >= hi[a]) { n[a] = 0; continue; }
            real t = (p - comp(S, a)) / comp(d, a);
            if (t <= tprev) { next[a] += step[a]; a--; continue; }
            if (t < tbest) { tbest = t; abest = a; }
        }
        if (tbest > tprev) {
            real tm = (tprev + tbest) / 2;
            int idx[3];
            for (int a = 0; a < 3; a++) {
                if (fixed_idx[a] >= 0) { idx[a] = fixed_idx[a]; continue; }
                int i = (int)floor((double)(comp(S, a) + tm * comp(d, a)));
                if (i < lo[a]) i = lo[a];
                if (i > hi[a] - 1) i = hi[a] - 1;
                idx[a] = i;
            }
            size_t off = ((size_t)idx[0] * dims[1] + idx[1]) * dims[2] + idx[2];
            out[off] += (tbest - tprev) * len * w;
            tprev = tbest;
        }
        if (abest < 0) break;
        next[abest] += step[abest];
    }
}

int OR_SUFFIX(or_backproject_matched)(const ap_config_cb_project* cfg,
                                      const ap_proj_spec* specs, uint32_t nprojs,
                                      const ap_vol_tex_chunk_spec* chunk,
                                      const uint32_t* dims, real* vol,
                                      const real* mult, const real* sino,
                                      real alpha) {
    if (!cfg || !specs || !chunk || !dims || !vol || !sino) return AP_ERR_INVALID;
    fwd_geom g = make_fwd_geom(cfg, chunk);
    size_t nvox = (size_t)dims[0] * dims[1] * dims[2];
    real* tmp = (real*)calloc(nvox, sizeof(real));
    if (!tmp) return AP_ERR_INVALID;
    const uint32_t nu = specs[0].dspec.nu, nv = specs[0].dspec.nv;
    real ssw = (real)cfg->multiply_before_adding / (real)(cfg->ssu * cfg->ssv);
    for (uint32_t p = 0; p < nprojs; p++) {
        v3 S = to_vox(&g, vf(specs[p].sourcepos));
        for (uint32_t j = 0; j < nu; j++)
            for (uint32_t k = 0; k < nv; k++) {
                real y = sino[((size_t)p * nu + j) * nv + k];
                if (y == 0) continue;
                for (uint32_t su = 0; su < cfg->ssu; su++)
                    for (uint32_t sv = 0; sv < cfg->ssv; sv++) {
                        real ju = (real)j - (real)0.5 + ((real)su + (real)0.5) / (real)cfg->ssu;
                        real kv = (real)k - (real)0.5 + ((real)sv + (real)0.5) / (real)cfg->ssv;
                        v3 Q = to_vox(&g, det_point(&specs[p].dspec, ju, kv));
                        siddon_scatter(S, Q, g.lo, g.hi, dims, y * ssw * g.voxwid, tmp);
                    }
            }
    }
    for (size_t i = 0; i < nvox; i++)
        vol[i] += alpha * (mult ? mult[i] : 1) * tmp[i];
    free(tmp);
    return AP_OK;
}

/* ----------------------------------------------------------------------
 * FP32 port of the reference's own traversal (snap-to-boundary DDA of
 * CB_project_texture_chunked.cu:111-272, one launch per call, no atomics
 * needed on the CPU).  Used as the timed CPU baseline of the forward
 * projector and cross-checked against the FP64 Siddon above.
 * ---------------------------------------------------------------------- */
static float refdda_unitvox(const float v1[3], const float v2[3],
                            const int lo[3], const int hi[3],
                            const uint32_t dims[3], const real* vol) {
    float st[3], en[3];
    for (int a = 0; a < 3; a++) {
        st[a] = v1[a];
        en[a] = v2[a];
        if (st[a] < lo[a] && en[a] < lo[a]) return 0;
        if (st[a] > hi[a] && en[a] > hi[a]) return 0;
    }
    float pos[3] = {st[0], st[1], st[2]};
    float dv[3] = {en[0] - st[0], en[1] - st[1], en[2] - st[2]};
    int dir[3];
    for (int a = 0; a < 3; a++) dir[a] = dv[a] < 0 ? -1 : (dv[a] == 0 ? 0 : 1);
    float len = sqrtf(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
    float remaining = 1.f;
    /* jump to half a voxel before the chunk */
    float adv = 0;
    for (int a = 0; a < 3; a++) {
        if (dir[a] < 0) adv = fmaxf(adv, ((float)hi[a] + 0.5f - pos[a]) / dv[a]);
        if (dir[a] > 0) adv = fmaxf(adv, ((float)lo[a] - 0.5f - pos[a]) / dv[a]);
    }
    if (adv > 0) {
        for (int a = 0; a < 3; a++) pos[a] += adv * dv[a];
        remaining -= adv;
    }
    float acc = 0;
    while (remaining > 0) {
        float nb[3], steps[3];
        int stop = 0;
        for (int a = 0; a < 3; a++) {
            nb[a] = dir[a] <= 0 ? ceilf(pos[a]) - 1 : floorf(pos[a]) + 1;
            if (dir[a] <= 0 && nb[a] < lo[a]) stop = 1;
            if (dir[a] > 0 && nb[a] > hi[a]) stop = 1;
            steps[a] = dir[a] == 0 ? 1.f : (nb[a] - pos[a]) / dv[a];
        }
        if (stop) break;
        int snap = 0;
        if (steps[0] <= steps[1]) snap = steps[0] <= steps[2] ? 0 : 2;
        else snap = steps[1] <= steps[2] ? 1 : 2;
        float take = fminf(steps[snap], remaining);
        int idx[3];
        for (int a = 0; a < 3; a++) idx[a] = (int)nb[a] - (1 + dir[a]) / 2;
        for (int a = 0; a < 3; a++) pos[a] = a == snap ? nb[a] : pos[a] + take * dv[a];
        if (idx[0] >= lo[0] && idx[0] < hi[0] && idx[1] >= lo[1] && idx[1] < hi[1] &&
            idx[2] >= lo[2] && idx[2] < hi[2]) {
            size_t off = ((size_t)idx[0] * dims[1] + idx[1]) * dims[2] + idx[2];
            acc += take * (float)vol[off];
        }
        remaining -= take;
    }
    return acc * len;
}

int OR_SUFFIX(or_project_voxel_refdda)(const ap_config_cb_project* cfg,
                                       const ap_proj_spec* specs, uint32_t nprojs,
                                       const ap_vol_tex_chunk_spec* chunk,
                                       const uint32_t* dims, const real* vol,
                                       real* sino, int nthreads) {
    if (!cfg || !specs || !chunk || !dims || !vol || !sino) return AP_ERR_INVALID;
    set_threads(nthreads);
    int lo[3] = {(int)chunk->offsets.x, (int)chunk->offsets.y, (int)chunk->offsets.z};
    int hi[3] = {lo[0] + (int)chunk->shape.x, lo[1] + (int)chunk->shape.y,
                 lo[2] + (int)chunk->shape.z};
    const float vw = chunk->voxwid, svf = cfg->shrink_volume_factor;
    const uint32_t nu = specs[0].dspec.nu, nv = specs[0].dspec.nv;
    long long total = (long long)nprojs * nu;
#pragma omp parallel for schedule(dynamic, 4)
    for (long long pj = 0; pj < total; pj++) {
        uint32_t p = (uint32_t)(pj / nu), j = (uint32_t)(pj % nu);
        const ap_proj_spec* ps = &specs[p];
        const ap_det_spec* ds = &ps->dspec;
        float s[3] = {ps->sourcepos.x * svf / vw - chunk->texpos.x / vw,
                      ps->sourcepos.y * svf / vw - chunk->texpos.y / vw,
                      ps->sourcepos.z * svf / vw - chunk->texpos.z / vw};
        for (uint32_t k = 0; k < nv; k++) {
            float tot = 0;
            for (uint32_t su = 0; su < cfg->ssu; su++) {
                float acc = 0;
                for (uint32_t sv = 0; sv < cfg->ssv; sv++) {
                    float xf = (float)j - 0.5f + (1.f / cfg->ssu) * (su + 0.5f);
                    float yf = (float)k - 0.5f + (1.f / cfg->ssv) * (sv + 0.5f);
                    float xn = xf - (ds->nu - 1) / 2.f, yn = yf - (ds->nv - 1) / 2.f;
                    float q[3] = {
                        ((ds->pos.x + ds->u.x * xn + ds->v.x * yn) * svf) / vw - chunk->texpos.x / vw,
                        ((ds->pos.y + ds->u.y * xn + ds->v.y * yn) * svf) / vw - chunk->texpos.y / vw,
                        ((ds->pos.z + ds->u.z * xn + ds->v.z * yn) * svf) / vw - chunk->texpos.z / vw};
                    acc += vw * refdda_unitvox(s, q, lo, hi, dims, vol) *
                           cfg->multiply_before_adding;
                }
                if (cfg->ssv > 1) acc /= (float)cfg->ssv;
                tot += acc;
            }
            if (cfg->ssu > 1) tot /= (float)cfg->ssu;
            sino[((size_t)p * nu + j) * nv + k] += tot;
        }
    }
    return AP_OK;
}

/* ======================================================================
 * Backprojection with the reference's semantics:
 * CB_backproject.cu:313-733 (_backproject_CB_getval) and :952-1027.
 * ====================================================================== */

/* get_projected_detector_coords (cuda/structs/geometry.py:75-107): detector
 * coordinates (in pixels, from the detector centre) of the central
 * projection of `x`. */
static void det_coords(const ap_proj_spec* ps, v3 x, real* cu, real* cv) {
    v3 S = vf(ps->sourcepos), D = vf(ps->dspec.pos);
    v3 u = vf(ps->dspec.u), vv = vf(ps->dspec.v);
    v3 n = cross(u, vv);
    real lam = dot(sub(D, S), n) / dot(sub(x, S), n);
    v3 p = sub(add(S, scl(sub(x, S), lam)), D);
    *cu = dot(u, p) / dot(u, u);
    *cv = dot(vv, p) / dot(vv, vv);
}

/* Length of the ray from S through P (infinite beyond P, starting at S)
 * inside the cube of side w centred on C (CB_backproject.cu:96-131). */
static real ray_cube_len(v3 S, v3 P, v3 C, real w) {
    v3 d = sub(P, S);
    real t0 = 0, t1 = INFINITY;
    for (int a = 0; a < 3; a++) {
        real s = comp(S, a), da = comp(d, a), c = comp(C, a);
        if (da == 0) {
            if (fabs((double)(s - c)) > (double)w / 2) return 0;
            continue;
        }
        real ta = (c - w / 2 - s) / da, tb = (c + w / 2 - s) / da;
        t0 = rmax(t0, rmin(ta, tb));
        t1 = rmin(t1, rmax(ta, tb));
    }
    return t1 > t0 ? (t1 - t0) * nrm(d) : 0;
}

static real powint(real b, int e) {
    real r = 1;
    if (e < 0) for (int i = 0; i < -e; i++) r /= b;
    else for (int i = 0; i < e; i++) r *= b;
    return r;
}

/* guarded detector read (zeros outside), CB_backproject.cu:153-170 */
static real det_read(const real* img, uint32_t nu, uint32_t nv, long j, long k) {
    if (j < 0 || k < 0 || j >= (long)nu || k >= (long)nv) return 0;
    return img[(size_t)j * nv + k];
}

/* bilinear (CB_backproject.cu:221-254) and nearest (:256-271) sampling at
 * fractional pixel index (fj, fk) measured from the centre of pixel (0,0) */
static real det_bilinear(const real* img, uint32_t nu, uint32_t nv, real fj, real fk) {
    if (fj <= -1 || fk <= -1 || fj >= (real)nu || fk >= (real)nv) return 0;
    long j0 = (long)floor((double)fj), k0 = (long)floor((double)fk);
    real a = fj - (real)j0, b = fk - (real)k0;
    real v00 = det_read(img, nu, nv, j0, k0), v01 = det_read(img, nu, nv, j0, k0 + 1);
    real v10 = det_read(img, nu, nv, j0 + 1, k0), v11 = det_read(img, nu, nv, j0 + 1, k0 + 1);
    real t1 = v00 * (1 - a) + v10 * a;
    real t2 = v01 * (1 - a) + v11 * a;
    return t1 * (1 - b) + t2 * b;
}

static real det_nearest(const real* img, uint32_t nu, uint32_t nv, real fj, real fk) {
    return det_read(img, nu, nv, (long)floor((double)(fj + (real)0.5)),
                    (long)floor((double)(fk + (real)0.5)));
}

/* Weighting presets of the direct ray-casting branch (CB_backproject.cu:
 * 528-600): the source moves on a cylinder of radius R about the z axis;
 * ell = |horizontal part of vp - S| / R, c = clamp(|Z / R / ell|, 0, 1),
 * d = c - ell, vy and vx the components of vp across and along the source
 * direction (units R^2).  Presets other than 1-4 and 6 leave mult alone. */
static real preset_mult(int preset, real R, v3 S, v3 vp, real Z, real mult) {
    v3 r = sub(vp, S);
    real ell = (real)sqrt((double)(r.x * r.x + r.y * r.y)) / R;
    real zr = Z / R;
    real c = rmax(0, rmin(1, (real)fabs((double)(zr / ell))));
    real d = c - ell;
    real vy = (S.x * vp.y - S.y * vp.x) / R / R;
    real vx = (S.x * vp.x + S.y * vp.y) / R / R;
    real q = d / c, f = (1 - q * q) / (1 + q * q);
    switch (preset) {
    case 1: return mult / (c * c * (c * c + d * d));
    case 2: return mult * (-2 * f * f * (-vy) * (1 + vx / zr) / (ell * ell));
    case 3: return mult / (zr * c * c);
    case 4: return mult * ((vy / zr) / zr);
    case 6: return mult * (-2 * f * f / (1 - q) * (-vy) * (1 + vx / zr) / zr);
    default: return mult;
    }
}

/* s(a, b; c) of the GBC paper's appendix (CB_backproject.cu:274-311) */
static real cossq_soft(real a, real b, real c) {
    real ac = (real)fabs((double)c);
    if (ac > a) return 0;
    if (ac < a - b) return 1;
    real g = 1 / ((real)sin((double)(a - b)) - (real)sin((double)a));
    real x = g * (real)sin((double)ac) - g * (real)sin((double)a);
    return rmax(0, rmin(1, 3 * x * x - 2 * x * x * x));
}

/* GBC weight of the source S for the point x, cylinder of radius R about the
 * z axis (CB_backproject.cu:736-790, eq. (23) of the GBC paper) */
static real gbc_weight(real R, real va, real vb, real ha, real hb, v3 S, v3 x) {
    v3 r = sub(x, S);
    real ln = (real)sqrt((double)(S.x * S.x + S.y * S.y));
    real Z = (r.x * -S.x + r.y * -S.y) / ln;
    real ell = (real)sqrt((double)(r.x * r.x + r.y * r.y)) / R;
    real c = rmax(0, rmin(1, (real)fabs((double)(Z / R / ell))));
    real d = c - ell;
    real th = (real)atan2((double)(r.z / R), (double)ell);
    real ce = (real)cos((double)th);
    real den = c * c + d * d;
    real w = ce * ce * ce * c / (den > (real)1e-9 ? den : (real)1e-9);
    return w * cossq_soft(va, vb, th) * cossq_soft(ha, hb, (real)acos((double)c));
}

real OR_SUFFIX(or_gbc_weight)(const ap_config_cb_backproject* cfg, const double* S3, const double* x3) {
    return gbc_weight((real)cfg->Z_on_val_factor_val, cfg->soft_vartheta_v, cfg->soft_vartheta_v_softspan,
                      cfg->soft_vartheta_h, cfg->soft_vartheta_h_softspan, mk((real)S3[0], (real)S3[1], (real)S3[2]),
                      mk((real)x3[0], (real)x3[1], (real)x3[2]));
}

/* contribution of one view to one (sub-)voxel at vox_pos */
static real bp_view(const ap_config_cb_backproject* cfg, const ap_proj_spec* ps,
                    const real* img, v3 vp, real w) {
    const ap_det_spec* ds = &ps->dspec;
    const uint32_t nu = ds->nu, nv = ds->nv;
    v3 S = vf(ps->sourcepos), D = vf(ds->pos);
    real cu, cv;
    det_coords(ps, vp, &cu, &cv);
    real val = 0;
    if (cfg->exact_line_integ_backproj) {
        real mnu = cu, mnv = cv, mxu = cu, mxv = cv;
        for (int c = 0; c < 8; c++) {
            v3 o = mk((c & 4) ? w / 2 : -w / 2, (c & 2) ? w / 2 : -w / 2, (c & 1) ? w / 2 : -w / 2);
            real a, b;
            det_coords(ps, add(vp, o), &a, &b);
            mnu = rmin(mnu, a); mxu = rmax(mxu, a);
            mnv = rmin(mnv, b); mxv = rmax(mxv, b);
        }
        long j0 = (long)floor((double)(mnu + (real)nu / 2)); if (j0 < 0) j0 = 0;
        long k0 = (long)floor((double)(mnv + (real)nv / 2)); if (k0 < 0) k0 = 0;
        long j1 = (long)ceil((double)(mxu + (real)nu / 2)); if (j1 > (long)nu - 1) j1 = nu - 1;
        long k1 = (long)ceil((double)(mxv + (real)nv / 2)); if (k1 > (long)nv - 1) k1 = nv - 1;
        for (long j = j0; j <= j1; j++)
            for (long k = k0; k <= k1; k++) {
                v3 P = det_point(ds, (real)j, (real)k);
                real len = ray_cube_len(S, P, vp, w);
                if (len != 0) {
                    real y = cfg->backproject_all_1s ? 1 : img[(size_t)j * nv + k];
                    val += y * len;
                }
            }
        if (cfg->Z_on_L_factor_power != 0) {
            v3 L = sub(D, S);
            val *= powint(dot(sub(vp, S), L) / dot(L, L), cfg->Z_on_L_factor_power);
        }
        if (cfg->Z_on_val_factor_power != 0) {
            v3 L = sub(D, S);
            val *= powint(dot(scl(sub(vp, S), 1 / (real)cfg->Z_on_val_factor_val), L) / nrm(L),
                          cfg->Z_on_val_factor_power);
        }
    } else {
        real fj = cu + (real)(nu - 1) / 2, fk = cv + (real)(nv - 1) / 2;
        real mult;
        if (cfg->approx_mode_direct_raycast) {
            mult = 1;
            if (cfg->Z_on_L_factor_power != 0 || cfg->Z_on_val_factor_power != 0 ||
                cfg->backproj_weighting_preset != 0) {
                v3 L = sub(D, S);
                real Z = dot(sub(vp, S), L) / nrm(L);
                if (cfg->Z_on_L_factor_power != 0)
                    mult *= powint(Z / nrm(L), cfg->Z_on_L_factor_power);
                if (cfg->Z_on_val_factor_power != 0)
                    mult *= powint(Z / (real)cfg->Z_on_val_factor_val, cfg->Z_on_val_factor_power);
                if (cfg->backproj_weighting_preset != 0)
                    mult = preset_mult(cfg->backproj_weighting_preset, (real)cfg->Z_on_val_factor_val, S, vp, Z,
                                       mult);
            }
        } else {
            v3 P = add(D, add(scl(vf(ds->u), cu), scl(vf(ds->v), cv)));
            real mag = nrm(sub(P, S)) / nrm(sub(vp, S)) * w;
            mult = mag * mag / (nrm(vf(ds->u)) * nrm(vf(ds->v)));
            if (cfg->Z_on_L_factor_power != 0) {
                v3 L = sub(D, S);
                mult *= powint(dot(sub(vp, S), L) / dot(L, L), cfg->Z_on_L_factor_power);
            }
            if (cfg->Z_on_val_factor_power != 0) {
                v3 L = sub(D, S);
                mult *= powint(dot(scl(sub(vp, S), 1 / (real)cfg->Z_on_val_factor_val), L) / nrm(L),
                               cfg->Z_on_val_factor_power);
            }
            mult *= w;
        }
        if (mult != 0) {
            real y = cfg->approx_mode_linear_interpolate ? det_bilinear(img, nu, nv, fj, fk)
                                                         : det_nearest(img, nu, nv, fj, fk);
            val = y * mult;
        }
    }
    real sep = cfg->soften_edge_proportion;
    if (sep > 0 && val != 0) {
        real ud = fabs((double)cu) / ((real)nu / 2), vd = fabs((double)cv) / ((real)nv / 2);
        ud = rmax(0, rmin(1, (1 - ud) / sep));
        vd = rmax(0, rmin(1, (1 - vd) / sep));
        ud = rmax(0, rmin(1, (1 + cos(3.14159265358979323846 * (1 - ud))) / 2));
        vd = rmax(0, rmin(1, (1 + cos(3.14159265358979323846 * (1 - vd))) / 2));
        val *= ud * vd;
    }
    return val;
}

/* vol[i] += mult[i] * BP(sino)[i] for every voxel of the chunk */
int OR_SUFFIX(or_backproject)(const ap_config_cb_backproject* cfg,
                              const ap_proj_spec* specs, uint32_t nprojs,
                              const ap_vol_tex_chunk_spec* chunk,
                              const uint32_t* dims, real* vol, const real* mult,
                              const real* sino, int nthreads) {
    if (!cfg || !specs || !chunk || !dims || !vol || !sino) return AP_ERR_INVALID;
    if (cfg->backproj_weighting_preset == 7) return AP_ERR_UNSUPPORTED;   /* or_backproject_gbc */
    if (cfg->ssx == 0 || cfg->ssy == 0 || cfg->ssz == 0) return AP_ERR_INVALID;
    set_threads(nthreads);
    const real w = chunk->voxwid;
    const v3 tp = vf(chunk->texpos);
    const uint32_t nu = specs[0].dspec.nu, nv = specs[0].dspec.nv;
    const uint32_t cx = chunk->shape.x, cy = chunk->shape.y, cz = chunk->shape.z;
    long long total = (long long)cx * cy;
#pragma omp parallel for schedule(dynamic, 4)
    for (long long ij = 0; ij < total; ij++) {
        uint32_t i = (uint32_t)(ij / cy), j = (uint32_t)(ij % cy);
        for (uint32_t k = 0; k < cz; k++) {
            uint32_t gi = i + chunk->offsets.x, gj = j + chunk->offsets.y, gk = k + chunk->offsets.z;
            v3 c = add(tp, scl(mk((real)gi + (real)0.5, (real)gj + (real)0.5, (real)gk + (real)0.5), w));
            if (cfg->mask_to_cylinder &&
                c.x * c.x + c.y * c.y > (real)cfg->mask_to_cyl_rad * (real)cfg->mask_to_cyl_rad)
                continue;
            real ax = 0;
            for (uint32_t si = 0; si < cfg->ssx; si++) {
                real ay = 0;
                for (uint32_t sj = 0; sj < cfg->ssy; sj++) {
                    real az = 0;
                    for (uint32_t sk = 0; sk < cfg->ssz; sk++) {
                        v3 vp = add(c, scl(mk(-(real)0.5 + ((real)0.5 + si) / cfg->ssx,
                                              -(real)0.5 + ((real)0.5 + sj) / cfg->ssy,
                                              -(real)0.5 + ((real)0.5 + sk) / cfg->ssz), w));
                        real acc = 0;
                        for (uint32_t p = 0; p < nprojs; p++)
                            acc += bp_view(cfg, &specs[p], sino + (size_t)p * nu * nv, vp, w);
                        az += acc;
                    }
                    ay += az / cfg->ssz;
                }
                ax += ay / cfg->ssy;
            }
            ax /= cfg->ssx;
            size_t off = ((size_t)gi * dims[1] + gj) * dims[2] + gk;
            real m = (cfg->apply_summand_multiplier && mult) ? mult[off] : 1;
            vol[off] += m * ax;
        }
    }
    return AP_OK;
}

/* GBC backprojection (backproject_CB_voltex_GBC_cyl, CB_backproject.cu:
 * 795-890, 1033-1108): vol += mult * mean_ss sum_views w * detector value,
 * sec += mult * mean_ss sum_views w, with w = gbc_weight and the detector
 * read bilinearly or at the nearest pixel. */
int OR_SUFFIX(or_backproject_gbc)(const ap_config_cb_backproject* cfg,
                                  const ap_proj_spec* specs, uint32_t nprojs,
                                  const ap_vol_tex_chunk_spec* chunk,
                                  const uint32_t* dims, real* vol, real* sec, const real* mult,
                                  const real* sino, int nthreads) {
    if (!cfg || !specs || !chunk || !dims || !vol || !sec || !sino) return AP_ERR_INVALID;
    if (cfg->ssx == 0 || cfg->ssy == 0 || cfg->ssz == 0) return AP_ERR_INVALID;
    set_threads(nthreads);
    const real w = chunk->voxwid, R = cfg->Z_on_val_factor_val;
    const v3 tp = vf(chunk->texpos);
    const uint32_t nu = specs[0].dspec.nu, nv = specs[0].dspec.nv;
    const uint32_t cx = chunk->shape.x, cy = chunk->shape.y, cz = chunk->shape.z;
    long long total = (long long)cx * cy;
#pragma omp parallel for schedule(dynamic, 4)
    for (long long ij = 0; ij < total; ij++) {
        uint32_t i = (uint32_t)(ij / cy), j = (uint32_t)(ij % cy);
        for (uint32_t k = 0; k < cz; k++) {
            uint32_t gi = i + chunk->offsets.x, gj = j + chunk->offsets.y, gk = k + chunk->offsets.z;
            v3 c = add(tp, scl(mk((real)gi + (real)0.5, (real)gj + (real)0.5, (real)gk + (real)0.5), w));
            real vx = 0, wx = 0;
            for (uint32_t si = 0; si < cfg->ssx; si++) {
                real vy = 0, wy = 0;
                for (uint32_t sj = 0; sj < cfg->ssy; sj++) {
                    real vz = 0, wz = 0;
                    for (uint32_t sk = 0; sk < cfg->ssz; sk++) {
                        v3 vp = add(c, scl(mk(-(real)0.5 + ((real)0.5 + si) / cfg->ssx,
                                              -(real)0.5 + ((real)0.5 + sj) / cfg->ssy,
                                              -(real)0.5 + ((real)0.5 + sk) / cfg->ssz), w));
                        for (uint32_t p = 0; p < nprojs; p++) {
                            const ap_proj_spec* ps = &specs[p];
                            real g = gbc_weight(R, cfg->soft_vartheta_v, cfg->soft_vartheta_v_softspan,
                                                cfg->soft_vartheta_h, cfg->soft_vartheta_h_softspan,
                                                vf(ps->sourcepos), vp);
                            wz += g;
                            if (g == 0) continue;
                            real cu, cv;
                            det_coords(ps, vp, &cu, &cv);
                            const real* img = sino + (size_t)p * nu * nv;
                            real fj = cu + (real)(nu - 1) / 2, fk = cv + (real)(nv - 1) / 2;
                            real y = cfg->approx_mode_linear_interpolate ? det_bilinear(img, nu, nv, fj, fk)
                                                                         : det_nearest(img, nu, nv, fj, fk);
                            vz += y * g;
                        }
                    }
                    vy += vz / cfg->ssz;
                    wy += wz / cfg->ssz;
                }
                vx += vy / cfg->ssy;
                wx += wy / cfg->ssy;
            }
            size_t off = ((size_t)gi * dims[1] + gj) * dims[2] + gk;
            real m = (cfg->apply_summand_multiplier && mult) ? mult[off] : 1;
            vol[off] += m * (vx / cfg->ssx);
            sec[off] += m * (wx / cfg->ssx);
        }
    }
    return AP_OK;
}

/* Integral of the GBC weight over the source cylinder (R = 1) for points
 * (r_i, 0, 0) (get_GBC_weight_integ, CB_backproject.cu:1111-1168): out[i] +=
 * 4 pi zmax / (nz nt) sum_{a<nt, b<nz} h_b w(t_a, z_b), t_a = pi a / nt,
 * z_b = zmax b / (nz - 1), zmax = tan(vartheta_v) (1 + r_{n-1}), h_0 = 1/2. */
int OR_SUFFIX(or_gbc_weight_integ)(const double* r, uint32_t n, double* out, uint32_t nz, uint32_t nt,
                                   const ap_config_cb_backproject* cfg) {
    if (!r || !out || !cfg || nz < 2 || nt < 1) return AP_ERR_INVALID;
    if (n == 0) return AP_OK;
    const double pi = 3.14159265358979323846;
    const double zmax = tan((double)cfg->soft_vartheta_v) * (1 + r[n - 1]);
#pragma omp parallel for schedule(dynamic, 1)
    for (long i = 0; i < (long)n; i++) {
        double acc = 0;
        for (uint32_t a = 0; a < nt; a++) {
            double t = pi * a / nt;
            for (uint32_t b = 0; b < nz; b++) {
                v3 S = mk((real)cos(t), (real)sin(t), (real)(zmax * b / (nz - 1)));
                real g = gbc_weight(1, cfg->soft_vartheta_v, cfg->soft_vartheta_v_softspan, cfg->soft_vartheta_h,
                                    cfg->soft_vartheta_h_softspan, S, mk((real)r[i], 0, 0));
                acc += (b == 0 ? 0.5 : 1.0) * (double)g;
            }
        }
        out[i] += 4 * pi * zmax / ((double)nz * nt) * acc;
    }
    return AP_OK;
}

/* ======================================================================
 * Ramp filter: direct linear convolution along axis 1 (u) with the taps
 * of ramp_kernels.py:27-94, optional FDK cosine pre-weight
 * (cb/ramp_filter.py:407-436).  Accumulated in long double.
 * ====================================================================== */
static long double ramp_tap(long n, long double p, int kernel) {
    const long double pi = 3.141592653589793238462643383279502884L;
    if (kernel == AP_RAMP_SINC) return 8.0L / (1.0L - 4.0L * (long double)n * n) / p;
    if (kernel == AP_RAMP_HILBERT) return n == 0 ? 0 : 2.0L / (long double)n;   /* :62-77 */
    if (kernel == AP_RAMP_HILBERT_HALF) return 2.0L / ((long double)n + 0.5L);  /* :80-94 */
    if (n == 0) return pi * pi / p;
    if (n % 2 == 0) return 0;
    return -4.0L / ((long double)n * n) / p;
}

int OR_SUFFIX(or_ramp_filter)(uint32_t nviews, uint32_t nu, uint32_t nv,
                              const real* in, real* out, double physpixwid,
                              int kernel, int fdk) {
    if (!in || !out || nu == 0) return AP_ERR_INVALID;
    long double p = physpixwid;
    for (uint32_t v = 0; v < nviews; v++)
        for (uint32_t k = 0; k < nv; k++)
            for (uint32_t j = 0; j < nu; j++) {
                long double acc = 0;
                for (uint32_t m = 0; m < nu; m++) {
                    long double x = in[((size_t)v * nu + m) * nv + k];
                    if (fdk) {
                        long double uu = p * ((long double)m - (long double)(nu - 1) / 2);
                        long double vv = p * ((long double)k - (long double)(nv - 1) / 2);
                        x /= sqrtl(1 + uu * uu + vv * vv);
                    }
                    acc += x * ramp_tap((long)j - (long)m, p, kernel);
                }
                out[((size_t)v * nu + j) * nv + k] = (real)acc;
            }
    return AP_OK;
}

/* ======================================================================
 * FORBILD analytic projector: exact intervals of the segment S->Q inside
 * each primitive (cuda/structs/cu/FORBILD_*.cu), last-wins compositing
 * (FORBILD_Shape.cu:474-540) by a sorted-breakpoint sweep.
 * ====================================================================== */
typedef struct { real a, b; } ival;   /* parametric interval in [0,1] */

static ival iv_empty(void) { ival r = {0, 0}; return r; }
static ival iv_cap(ival x, ival y) {
    ival r = {rmax(x.a, y.a), rmin(x.b, y.b)};
    if (r.b < r.a) r.b = r.a;
    return r;
}

/* |S + t d|^2 <= 1 in a frame where the quadric is the unit sphere */
static ival unit_sphere(v3 S, v3 d) {
    real A = dot(d, d), B = dot(S, d), C = dot(S, S) - 1;
    if (A == 0) return iv_empty();
    real disc = B * B - A * C;
    if (disc <= 0) return iv_empty();
    real sq = (real)sqrt((double)disc);
    /* numerically stable pair of roots */
    real q = B > 0 ? -(B + sq) : -(B - sq);
    real r1 = q / A, r2 = (q != 0) ? C / q : -r1;
    ival r = {rmin(r1, r2), rmax(r1, r2)};
    ival unit = {0, 1};
    return iv_cap(r, unit);
}

/* x^2 + y^2 <= 1 and |z| <= l/2 in the cylinder frame */
static ival unit_cyl(v3 S, v3 d, real l) {
    ival r = {0, 1};
    if (d.z == 0) {
        if (fabs((double)S.z) > (double)l / 2) return iv_empty();
    } else {
        real ta = (l / 2 - S.z) / d.z, tb = (-l / 2 - S.z) / d.z;
        ival zz = {rmin(ta, tb), rmax(ta, tb)};
        r = iv_cap(r, zz);
    }
    real A = d.x * d.x + d.y * d.y, B = S.x * d.x + S.y * d.y;
    real C = S.x * S.x + S.y * S.y - 1;
    if (A == 0) {
        if (C > 0) return iv_empty();
        return r;
    }
    real disc = B * B - A * C;
    if (disc <= 0) return iv_empty();
    real sq = (real)sqrt((double)disc);
    real q = B > 0 ? -(B + sq) : -(B - sq);
    real r1 = q / A, r2 = (q != 0) ? C / q : -r1;
    ival rr = {rmin(r1, r2), rmax(r1, r2)};
    return iv_cap(r, rr);
}

static v3 unit3(ap_fvec3 a) { v3 r = vf(a); return scl(r, 1 / nrm(r)); }

/* Truncated cone in its frame (axis z): |z| <= l/2 and
 * x^2 + y^2 <= rad(z)^2, rad(z) = r1 + (z/l + 1/2)(r2 - r1)
 * (cuda/structs/cu/FORBILD_Cone_z.cu:1-13).  Breakpoints are the slab ends
 * and the real roots of x^2 + y^2 - rad^2 along the ray; every piece between
 * consecutive breakpoints is inside or outside as a whole, decided by the
 * membership test at its midpoint. */
static int in_cone_frame(v3 p, real l, real r1, real r2) {
    if (fabs((double)p.z) > (double)l / 2) return 0;
    real rad = r1 + (p.z / l + (real)0.5) * (r2 - r1);
    return rad >= 0 && p.x * p.x + p.y * p.y <= rad * rad;
}

static ival unit_cone(v3 S, v3 d, real l, real r1, real r2) {
    real br[6];
    int nb = 0;
    br[nb++] = 0;
    br[nb++] = 1;
    if (d.z != 0) {
        br[nb++] = (l / 2 - S.z) / d.z;
        br[nb++] = (-l / 2 - S.z) / d.z;
    }
    /* rad(z) = ca + cb z along S + t d */
    real cb = (r2 - r1) / l, ca = (r1 + r2) / 2;
    real g0 = ca + cb * S.z, gd = cb * d.z;
    real A = d.x * d.x + d.y * d.y - gd * gd;
    real B = 2 * (S.x * d.x + S.y * d.y - g0 * gd);
    real C = S.x * S.x + S.y * S.y - g0 * g0;
    if (A != 0) {
        real disc = B * B - 4 * A * C;
        if (disc > 0) {
            real sq = (real)sqrt((double)disc);
            br[nb++] = (-B - sq) / (2 * A);
            br[nb++] = (-B + sq) / (2 * A);
        }
    } else if (B != 0) {
        br[nb++] = -C / B;
    }
    /* sort the breakpoints inside [0,1] */
    real pts[6];
    int np = 0;
    for (int i = 0; i < nb; i++)
        if (br[i] >= 0 && br[i] <= 1) pts[np++] = br[i];
    for (int i = 1; i < np; i++)
        for (int j = i; j > 0 && pts[j] < pts[j - 1]; j--) { real t = pts[j]; pts[j] = pts[j - 1]; pts[j - 1] = t; }
    ival r = iv_empty();
    int found = 0;
    for (int i = 0; i + 1 < np; i++) {
        if (!(pts[i + 1] > pts[i])) continue;
        real mid = (pts[i] + pts[i + 1]) / 2;
        if (in_cone_frame(add(S, scl(d, mid)), l, r1, r2)) {
            if (!found) { r.a = pts[i]; found = 1; }
            r.b = pts[i + 1];
        }
    }
    return r;
}

static ival clip_plane(const ap_forbild_clip* c, v3 S, v3 d) {
    ival full = {0, 1};
    if (c->mode == 0) return full;
    v3 ax = vf(c->axis);
    real s0 = dot(S, ax) - c->dotprod, sd = dot(d, ax);
    /* mode 1 keeps ax.r <= dotprod; mode 2 keeps ax.r > dotprod */
    real sgn = c->mode == 1 ? 1 : -1;
    s0 *= sgn; sd *= sgn;          /* keep s0 + t sd <= 0 */
    if (sd == 0) return s0 <= 0 ? full : iv_empty();
    real t = -s0 / sd;
    ival r;
    if (sd > 0) { r.a = 0; r.b = t; } else { r.a = t; r.b = 1; }
    return iv_cap(r, full);
}

static ival shape_interval(const ap_forbild_shape* s, v3 S, v3 Q) {
    v3 d = sub(Q, S);
    ival iv = {0, 1};
    iv = iv_cap(iv, clip_plane(&s->cas1, S, d));
    iv = iv_cap(iv, clip_plane(&s->cas2, S, d));
    iv = iv_cap(iv, clip_plane(&s->cas3, S, d));
    if (!(iv.b > iv.a)) return iv_empty();
    v3 c = mk(s->x, s->y, s->z);
    v3 Sr = sub(S, c);
    ival prim;
    switch (s->type) {
    case 1: { /* Box: full side lengths, axis aligned */
        prim.a = 0; prim.b = 1;
        real hw[3] = {(real)s->dx / 2, (real)s->dy / 2, (real)s->dz / 2};
        for (int a = 0; a < 3; a++) {
            real sa = comp(Sr, a), da = comp(d, a);
            if (da == 0) {
                if (fabs((double)sa) > (double)hw[a]) return iv_empty();
                continue;
            }
            real ta = (-hw[a] - sa) / da, tb = (hw[a] - sa) / da;
            ival q = {rmin(ta, tb), rmax(ta, tb)};
            prim = iv_cap(prim, q);
        }
        break;
    }
    case 6: case 7: case 8: case 9: { /* circular cylinder, axis generic/x/y/z */
        v3 ax = s->type == 6 ? unit3(s->axis)
              : s->type == 7 ? mk(1, 0, 0) : s->type == 8 ? mk(0, 1, 0) : mk(0, 0, 1);
        /* orthonormal frame (e1, e2, ax) */
        v3 t = fabs((double)ax.x) < 0.9 ? mk(1, 0, 0) : mk(0, 1, 0);
        v3 e1 = cross(ax, t); e1 = scl(e1, 1 / nrm(e1));
        v3 e2 = cross(ax, e1);
        real r = s->r;
        v3 Sl = mk(dot(Sr, e1) / r, dot(Sr, e2) / r, dot(Sr, ax));
        v3 dl = mk(dot(d, e1) / r, dot(d, e2) / r, dot(d, ax));
        prim = unit_cyl(Sl, dl, s->l);
        break;
    }
    case 2: case 3: case 4: case 5: { /* cone, axis generic/x/y/z */
        v3 ax = s->type == 2 ? unit3(s->axis)
              : s->type == 3 ? mk(1, 0, 0) : s->type == 4 ? mk(0, 1, 0) : mk(0, 0, 1);
        v3 t = fabs((double)ax.x) < 0.9 ? mk(1, 0, 0) : mk(0, 1, 0);
        v3 e1 = cross(ax, t); e1 = scl(e1, 1 / nrm(e1));
        v3 e2 = cross(ax, e1);
        v3 Sl = mk(dot(Sr, e1), dot(Sr, e2), dot(Sr, ax));
        v3 dl = mk(dot(d, e1), dot(d, e2), dot(d, ax));
        prim = unit_cone(Sl, dl, s->l, s->r1, s->r2);
        break;
    }
    case 10: { /* axis-aligned ellipsoid, semi-axes dx,dy,dz */
        v3 Sl = mk(Sr.x / s->dx, Sr.y / s->dy, Sr.z / s->dz);
        v3 dl = mk(d.x / s->dx, d.y / s->dy, d.z / s->dz);
        prim = unit_sphere(Sl, dl);
        break;
    }
    case 11: { /* free ellipsoid: semi-axes along normalised a_x,a_y,a_z */
        v3 ex = unit3(s->a_x), ey = unit3(s->a_y), ez = unit3(s->a_z);
        v3 Sl = mk(dot(Sr, ex) / s->dx, dot(Sr, ey) / s->dy, dot(Sr, ez) / s->dz);
        v3 dl = mk(dot(d, ex) / s->dx, dot(d, ey) / s->dy, dot(d, ez) / s->dz);
        prim = unit_sphere(Sl, dl);
        break;
    }
    case 12: case 14: case 15: case 16: { /* elliptic cylinder */
        v3 ex, ey, ez;
        if (s->type == 12) { ex = unit3(s->a_x); ey = unit3(s->a_y); ez = unit3(s->axis); }
        else if (s->type == 14) { ex = mk(0, 1, 0); ey = mk(0, 0, 1); ez = mk(1, 0, 0); }
        else if (s->type == 15) { ex = mk(0, 0, 1); ey = mk(1, 0, 0); ez = mk(0, 1, 0); }
        else { ex = mk(1, 0, 0); ey = mk(0, 1, 0); ez = mk(0, 0, 1); }
        v3 Sl = mk(dot(Sr, ex) / s->dx, dot(Sr, ey) / s->dy, dot(Sr, ez));
        v3 dl = mk(dot(d, ex) / s->dx, dot(d, ey) / s->dy, dot(d, ez));
        prim = unit_cyl(Sl, dl, s->l);
        break;
    }
    case 13: { /* sphere */
        real r = s->r;
        prim = unit_sphere(scl(Sr, 1 / r), scl(d, 1 / r));
        break;
    }
    default:
        return iv_empty();
    }
    iv = iv_cap(iv, prim);
    return iv.b > iv.a ? iv : iv_empty();
}

static int cmp_real(const void* a, const void* b) {
    real x = *(const real*)a, y = *(const real*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* last-wins (or additive) integral over [0,1] of the ray S->Q, times |Q-S| */
static real forbild_ray(const ap_forbild_shape* shapes, uint32_t K, int additive,
                        v3 S, v3 Q, ival* iv, real* pts) {
    real len = nrm(sub(Q, S));
    if (len == 0) return 0;
    uint32_t m = 0;
    for (uint32_t s = 0; s < K; s++) {
        iv[s] = shape_interval(&shapes[s], S, Q);
        if (iv[s].b > iv[s].a) { pts[m++] = iv[s].a; pts[m++] = iv[s].b; }
    }
    if (m == 0) return 0;
    real acc = 0;
    if (additive) {
        for (uint32_t s = 0; s < K; s++) acc += (iv[s].b - iv[s].a) * shapes[s].rho;
        return acc * len;
    }
    qsort(pts, m, sizeof(real), cmp_real);
    for (uint32_t e = 0; e + 1 < m; e++) {
        real a = pts[e], b = pts[e + 1];
        if (!(b > a)) continue;
        real mid = (a + b) / 2;
        for (int32_t s = (int32_t)K - 1; s >= 0; s--)
            if (iv[s].b > iv[s].a && iv[s].a <= mid && mid <= iv[s].b) {
                acc += (b - a) * shapes[s].rho;
                break;
            }
    }
    return acc * len;
}

int OR_SUFFIX(or_project_forbild)(const ap_config_cb_project_forbild* cfg,
                                  const ap_proj_spec* specs, uint32_t nprojs,
                                  const ap_forbild_shape* shapes, uint32_t K,
                                  real* sino, int nthreads) {
    if (!cfg || !specs || !sino || (K && !shapes)) return AP_ERR_INVALID;
    set_threads(nthreads);
    const uint32_t nu = specs[0].dspec.nu, nv = specs[0].dspec.nv;
    long long total = (long long)nprojs * nu;
    int err = 0;
#pragma omp parallel
    {
        ival* iv = (ival*)malloc(sizeof(ival) * (K + 1));
        real* pts = (real*)malloc(sizeof(real) * (2 * K + 2));
        if (!iv || !pts) err = 1;
#pragma omp for schedule(dynamic, 4)
        for (long long pj = 0; pj < total; pj++) {
            if (!iv || !pts) continue;
            uint32_t p = (uint32_t)(pj / nu), j = (uint32_t)(pj % nu);
            v3 S = vf(specs[p].sourcepos);
            for (uint32_t k = 0; k < nv; k++) {
                real tot = 0;
                for (uint32_t su = 0; su < cfg->ssu; su++) {
                    real acc = 0;
                    for (uint32_t sv = 0; sv < cfg->ssv; sv++) {
                        real ju = (real)j - (real)0.5 + ((real)su + (real)0.5) / cfg->ssu;
                        real kv = (real)k - (real)0.5 + ((real)sv + (real)0.5) / cfg->ssv;
                        v3 Q = det_point(&specs[p].dspec, ju, kv);
                        acc += forbild_ray(shapes, K, cfg->overlap_addition, S, Q, iv, pts);
                    }
                    tot += acc / cfg->ssv;
                }
                sino[((size_t)p * nu + j) * nv + k] += tot / cfg->ssu * cfg->multiply_before_adding;
            }
        }
        free(iv);
        free(pts);
    }
    return err ? AP_ERR_INVALID : AP_OK;
}

/* point membership of one primitive, clip planes included
 * (the _in_region functions of cuda/structs/cu/FORBILD_*.cu) */
static int in_shape(const ap_forbild_shape* s, v3 p) {
    const ap_forbild_clip* cs[3] = {&s->cas1, &s->cas2, &s->cas3};
    for (int i = 0; i < 3; i++) {
        if (cs[i]->mode == 0) continue;
        real dp = dot(p, vf(cs[i]->axis));
        if (cs[i]->mode == 1 && !(dp <= cs[i]->dotprod)) return 0;
        if (cs[i]->mode == 2 && !(dp > cs[i]->dotprod)) return 0;
    }
    v3 r = sub(p, mk(s->x, s->y, s->z));
    switch (s->type) {
    case 1:
        return fabs((double)r.x) <= s->dx / 2.0 && fabs((double)r.y) <= s->dy / 2.0 &&
               fabs((double)r.z) <= s->dz / 2.0;
    case 6: case 7: case 8: case 9: {
        v3 ax = s->type == 6 ? unit3(s->axis)
              : s->type == 7 ? mk(1, 0, 0) : s->type == 8 ? mk(0, 1, 0) : mk(0, 0, 1);
        real z = dot(r, ax);
        if (fabs((double)z) > s->l / 2.0) return 0;
        v3 rr = sub(r, scl(ax, z));
        return dot(rr, rr) <= (real)s->r * s->r;
    }
    case 2: case 3: case 4: case 5: {
        v3 ax = s->type == 2 ? unit3(s->axis)
              : s->type == 3 ? mk(1, 0, 0) : s->type == 4 ? mk(0, 1, 0) : mk(0, 0, 1);
        real z = dot(r, ax);
        v3 rr = sub(r, scl(ax, z));
        return in_cone_frame(mk(nrm(rr), 0, z), s->l, s->r1, s->r2);
    }
    case 10: {
        real a = r.x / s->dx, b = r.y / s->dy, c = r.z / s->dz;
        return a * a + b * b + c * c <= 1;
    }
    case 11: {
        real a = dot(r, unit3(s->a_x)) / s->dx, b = dot(r, unit3(s->a_y)) / s->dy,
             c = dot(r, unit3(s->a_z)) / s->dz;
        return a * a + b * b + c * c <= 1;
    }
    case 12: case 14: case 15: case 16: {
        v3 ex, ey, ez;
        if (s->type == 12) { ex = unit3(s->a_x); ey = unit3(s->a_y); ez = unit3(s->axis); }
        else if (s->type == 14) { ex = mk(0, 1, 0); ey = mk(0, 0, 1); ez = mk(1, 0, 0); }
        else if (s->type == 15) { ex = mk(0, 0, 1); ey = mk(1, 0, 0); ez = mk(0, 1, 0); }
        else { ex = mk(1, 0, 0); ey = mk(0, 1, 0); ez = mk(0, 0, 1); }
        if (fabs((double)dot(r, ez)) > s->l / 2.0) return 0;
        real a = dot(r, ex) / s->dx, b = dot(r, ey) / s->dy;
        return a * a + b * b <= 1;
    }
    case 13:
        return dot(r, r) <= (real)s->r * s->r;
    default:
        return 0;
    }
}

/* add_FORBILD_sampling (sample_FORBILD_volumes.cu:10-123): vol += mean over
 * ss^3 sub-points (offsets (o-(ss-1)/2)/ss voxels) of the last shape's rho */
int OR_SUFFIX(or_voxelise_forbild)(const ap_forbild_shape* shapes, uint32_t K,
                                   const ap_vol_tex_chunk_spec* chunk,
                                   const uint32_t* dims, real* vol, uint32_t ss,
                                   int nthreads) {
    if (!chunk || !dims || !vol || ss == 0 || (K && !shapes)) return AP_ERR_INVALID;
    set_threads(nthreads);
    const real w = chunk->voxwid;
    const v3 tp = vf(chunk->texpos);
    long long total = (long long)chunk->shape.x * chunk->shape.y;
#pragma omp parallel for schedule(dynamic, 4)
    for (long long ij = 0; ij < total; ij++) {
        uint32_t i = (uint32_t)(ij / chunk->shape.y) + chunk->offsets.x;
        uint32_t j = (uint32_t)(ij % chunk->shape.y) + chunk->offsets.y;
        for (uint32_t k = chunk->offsets.z; k < chunk->offsets.z + chunk->shape.z; k++) {
            v3 c = add(tp, scl(mk(i + (real)0.5, j + (real)0.5, k + (real)0.5), w));
            real tot = 0;
            for (uint32_t a = 0; a < ss; a++)
                for (uint32_t b = 0; b < ss; b++)
                    for (uint32_t e = 0; e < ss; e++) {
                        v3 p = add(c, scl(mk(((real)a - (ss - (real)1) / 2) / ss,
                                             ((real)b - (ss - (real)1) / 2) / ss,
                                             ((real)e - (ss - (real)1) / 2) / ss), w));
                        real val = 0;
                        for (uint32_t s = 0; s < K; s++)
                            if (in_shape(&shapes[s], p)) val = shapes[s].rho;
                        tot += val;
                    }
            vol[((size_t)i * dims[1] + j) * dims[2] + k] += tot / ((real)ss * ss * ss);
        }
    }
    return AP_OK;
}
