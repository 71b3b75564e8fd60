/*
 * aprict_hip.h -- C ABI of the MI355X cone-beam projection library
 * (libaprict_hip.so, hand-written HIP kernels for gfx950).
 *
 * Conventions shared by every entry point
 *   - All array pointers are DEVICE pointers; descriptors (ap_array3, the
 *     config structs) are HOST pointers read during the call.
 *   - `specs` is a device array of `num_projs` ap_proj_spec (one per view).
 *   - `stream` is a hipStream_t passed as void* (NULL = legacy stream).
 *   - Every call is asynchronous on `stream`, allocates no device memory,
 *     and returns AP_OK (0) or a negative AP_ERR_* code; the message of the
 *     last failure on the calling thread is ap_last_error().
 *   - No C++ exception crosses this boundary.
 *
 * Reference interfaces replaced (paths relative to aprict's src/aprict/):
 *   the POD structs mirror the kernel-argument structs of
 *   cuda/structs/{geometry,config_cb_project,config_cb_project_FORBILD,
 *   config_cb_backproject,FORBILD_volumes}.py field for field (same names,
 *   order, natural alignment), so a numpy structured array built with the
 *   reference dtypes can be passed unchanged.
 */
#ifndef APRICT_HIP_H
#define APRICT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define AP_OK                0
#define AP_ERR_INVALID      -1   /* bad argument / descriptor             */
#define AP_ERR_UNSUPPORTED  -2   /* valid but not implemented combination */
#define AP_ERR_HIP          -3   /* HIP runtime error (launch, ...)        */

/* ---- element types of ap_array3 ---------------------------------------- */
#define AP_F32   0
#define AP_F16   1
#define AP_BF16  2
#define AP_F64   3

/* ---- ramp filter kernels ------------------------------------------------ */
#define AP_RAMP_RAMLAK   0   /* ramp_kernels.py:43-59 */
#define AP_RAMP_SINC     1   /* ramp_kernels.py:17-40 (ramp_kernel) */
#define AP_RAMP_HILBERT  2   /* ramp_kernels.py:62-77 (hilbert_transform_kernel) */
#define AP_RAMP_HILBERT_HALF 3   /* ramp_kernels.py:80-94 (..._halfoffset) */

/* detector weighting modes of the FBP filter (ap_fbp_rotate_weight) */
#define AP_FBP_GBC   0   /* cone weight + letterbox (ramp_filter.py:437-449, 689-754) */
#define AP_FBP_FAN   1   /* cone weight, no letterbox (_hacky_fan_beam_mode)          */
#define AP_FBP_FDK   2   /* cosine weight, no letterbox (_hacky_fdk_mode, :425-436)   */

/* ---- reference kernel-argument structs --------------------------------- */
typedef struct { float x, y, z; } ap_fvec3;
typedef struct { uint32_t x, y, z; } ap_uivec3;
typedef struct { uint64_t x, y, z; } ap_ulvec3;

/* det_spec: cuda/structs/geometry.py:40-53 (48 bytes) */
typedef struct {
    ap_fvec3 pos;      /* detector centre                         */
    ap_fvec3 u;        /* vector between adjacent pixel centres    */
    ap_fvec3 v;        /*   along u (row) and v (column)           */
    uint32_t nu;
    uint32_t nv;
    uint8_t  type;     /* 0 = flat panel (the only supported type) */
} ap_det_spec;

/* proj_spec: cuda/structs/geometry.py:56-109 (60 bytes) */
typedef struct {
    ap_fvec3    sourcepos;
    ap_det_spec dspec;
} ap_proj_spec;

/* proj_chunk_spec_maxlen10: cuda/structs/geometry.py:112-123 (640 bytes).
 * The by-value argument of one reference backprojection launch
 * (CB_backproject.cu:952-965): up to 10 views, the device address of the
 * first view's projection data and its byte strides (view, u, v).  The
 * entry points here take the views as a device array instead and need no
 * 10-view chunks; the type is defined so that a caller holding these
 * structs can map them: specs = projspecs[0..numprojs), data at
 * projmemaddr with element strides projmemstrides / element size.        */
typedef struct {
    ap_proj_spec projspecs[10];
    uint8_t      numprojs;
    uint64_t     projmemaddr;
    ap_ulvec3    projmemstrides;
} ap_proj_chunk_spec_maxlen10;

/* vol_tex_chunk_spec: cuda/structs/geometry.py:175-186 (40 bytes).
 * texpos = corner of voxel (0,0,0) of the whole volume; offsets/shape
 * select the chunk (in voxels) that a call reads or writes.  Every entry
 * point addresses a volume (and a multiplier) as
 *   vol + (i*strides[0] + j*strides[1] + k*strides[2]) * itemsize
 * with whole-volume indices (i,j,k) and touches no element outside the
 * chunk, so only the chunk needs to be backed by memory: a caller streaming
 * a volume larger than device memory passes, for a brick buffer holding the
 * chunk, vol = brick - (offsets . strides) * itemsize and the brick's own
 * strides. */
typedef struct {
    float     voxwid;
    ap_uivec3 offsets;
    ap_uivec3 shape;
    ap_fvec3  texpos;
} ap_vol_tex_chunk_spec;

/* config_CB_project: cuda/structs/config_cb_project.py:27-39 (16 bytes) */
typedef struct {
    float    shrink_volume_factor;
    float    multiply_before_adding;
    uint32_t ssu;
    uint32_t ssv;
} ap_config_cb_project;

/* config_CB_project_FORBILD: cuda/structs/config_cb_project_FORBILD.py:28-40 (16 bytes) */
typedef struct {
    float    multiply_before_adding;
    uint8_t  overlap_addition;
    uint32_t ssu;
    uint32_t ssv;
} ap_config_cb_project_forbild;

/* config_CB_backproject: cuda/structs/config_cb_backproject.py:27-72 (60 bytes) */
typedef struct {
    uint8_t  exact_line_integ_backproj;
    uint8_t  backproject_all_1s;
    uint8_t  mask_to_cylinder;
    float    mask_to_cyl_rad;            /* physical units */
    uint8_t  apply_summand_multiplier;
    float    soften_edge_proportion;
    int32_t  Z_on_L_factor_power;
    int32_t  Z_on_val_factor_power;
    float    Z_on_val_factor_val;
    uint8_t  approx_mode_direct_raycast;
    uint8_t  approx_mode_linear_interpolate;
    uint8_t  backproj_weighting_preset;
    uint32_t ssx;
    uint32_t ssy;
    uint32_t ssz;
    float    soft_vartheta_v;
    float    soft_vartheta_v_softspan;
    float    soft_vartheta_h;
    float    soft_vartheta_h_softspan;
} ap_config_cb_backproject;

/* FORBILD_ClipAxisSpec: cuda/structs/FORBILD_volumes.py:502-514 (20 bytes).
 * mode 0 = unused, 1 = keep axis.r <= dotprod, 2 = keep axis.r > dotprod */
typedef struct {
    ap_fvec3 axis;
    float    dotprod;
    uint8_t  mode;
} ap_forbild_clip;

/* FORBILD_Shape: cuda/structs/FORBILD_volumes.py:537-607 (156 bytes).
 * type ids: FORBILD_volumes.py:517-534 (1 Box, 2-5 Cone[_x/_y/_z],
 * 6-9 Cylinder[_x/_y/_z], 10 Ellipsoid, 11 Ellipsoid_free, 12 Ellipt_Cyl,
 * 13 Sphere, 14-16 Ellipt_Cyl_x/_y/_z). */
typedef struct {
    uint8_t         type;
    float           rho;
    ap_forbild_clip cas1, cas2, cas3;
    float           x, y, z, dx, dy, dz, r, r1, r2, l;
    ap_fvec3        a_x, a_y, a_z, axis;
} ap_forbild_shape;

/* ---- this library's array descriptor (no reference counterpart) -------- */
/* A strided 3-D device array: volume [Nx][Ny][Nz] or sinogram [V][nu][nv].
 * strides are in ELEMENTS (torch convention), may be any non-negative
 * values (C order, Fortran order, views). */
typedef struct {
    int32_t  dtype;        /* AP_F32 / AP_F16 / AP_BF16 / AP_F64 */
    uint32_t shape[3];
    int64_t  strides[3];
} ap_array3;

/* ---- library ----------------------------------------------------------- */
const char* ap_last_error(void);
int32_t     ap_version(void);
/* sizeof of every ABI struct, for binding-side layout checks:
 * which = 0 det_spec, 1 proj_spec, 2 vol_tex_chunk_spec, 3 config_cb_project,
 * 4 config_cb_project_forbild, 5 config_cb_backproject, 6 forbild_clip,
 * 7 forbild_shape, 8 array3.  Returns -1 for an unknown id. */
int32_t     ap_struct_size(int32_t which);

/* ---- forward projection (voxel volume) ---------------------------------
 * Replaces the project_CB_texture_chunked launch in cb/flatpanel.py:997-1125
 * (kernel cuda/cusrc_templates/CB_project_texture_chunked.cu:327-412).
 * For every view p, pixel (j,k):
 *   sino[p][j][k] = base + multiply_before_adding *
 *                   mean over ssu x ssv sub-rays of  INT_{source->subpixel} f
 * with base = b[p][j][k] if b != NULL (fused residual) else sino[p][j][k]
 * (accumulate, like the reference's atomicAdd).  b has sino's shape and
 * dtype and its own strides (b_desc, required with b; b may be sino itself
 * with the same strides).  f is the piecewise-constant volume chunk `chunk`,
 * stored in `vol` (desc vol_desc: dtype F32/F16/BF16, shape = whole volume).
 * Sinogram dtype: F32 or F16.                                              */
int32_t ap_project_voxel(const ap_config_cb_project* cfg,
                         const ap_proj_spec* specs, uint32_t num_projs,
                         const ap_vol_tex_chunk_spec* chunk,
                         const ap_array3* vol_desc, const void* vol,
                         const ap_array3* sino_desc, void* sino,
                         const ap_array3* b_desc, const void* b, void* stream);

/* ---- backprojection (reference semantics) ------------------------------
 * Replaces the backproject_CB_voltex launches in cb/flatpanel.py:1128-1281
 * (kernel cuda/cusrc_templates/CB_backproject.cu:952-1027; exact branch
 * :408-477, fast branch :478-691, softening :694-713).  For every voxel i of
 * the chunk:  vol[i] += (multiplier ? multiplier[i] : 1) * BP_cfg(sino)[i].
 * backproj_weighting_preset 1-4 and 6 weight the direct ray-casting branch
 * (:528-600; R = Z_on_val_factor_val).  Preset 7 replaces the whole
 * backprojection by the GBC one for a cylindrical source locus of radius
 * R about the z axis (backproject_CB_voltex_GBC_cyl, :736-1108): vol gets
 * sum_views w * detector value, `secondary` (required, same dtype and shape
 * as vol) gets sum_views w, both times the multiplier.  `secondary` is
 * ignored for the other presets, as in the reference.                      */
int32_t ap_backproject(const ap_config_cb_backproject* cfg,
                       const ap_proj_spec* specs, uint32_t num_projs,
                       const ap_vol_tex_chunk_spec* chunk,
                       const ap_array3* vol_desc, void* vol,
                       const ap_array3* mult_desc, const void* multiplier,
                       const ap_array3* sino_desc, const void* sino,
                       const ap_array3* sec_desc, void* secondary,
                       void* stream);

/* ---- matched adjoint of ap_project_voxel --------------------------------
 * vol[i] += alpha * (multiplier ? multiplier[i] : 1) * (A^T sino)[i]
 * where A is exactly the operator of ap_project_voxel with the same cfg
 * (same Siddon weights, ssu/ssv averaging, shrink factor and
 * multiply_before_adding).  No reference kernel computes this transpose;
 * at ssu=ssv=1 it equals the exact backprojector of CB_backproject.cu:408-477
 * (ConfigBackproject(fast_approx=False) without mask/softening/Z-weights).
 *
 * `workspace` (device memory, 256-byte aligned, may be NULL) holds the
 * per-column prefix sums of the gather formulation that upright-detector
 * views take: ap_backproject_matched_workspace_size() bytes let it take all
 * of them in one pass; less takes them in view chunks that fit (each chunk
 * adds to vol); NULL or too little for one view leaves every view to the
 * general (brick) kernel.  Results agree to FP32 rounding in every case.
 * The workspace is scratch: calls that share one must be ordered.           */
int32_t ap_backproject_matched(const ap_config_cb_project* cfg,
                               const ap_proj_spec* specs, uint32_t num_projs,
                               const ap_vol_tex_chunk_spec* chunk,
                               const ap_array3* vol_desc, void* vol,
                               const ap_array3* mult_desc, const void* multiplier,
                               const ap_array3* sino_desc, const void* sino,
                               float alpha, void* workspace, uint64_t workspace_bytes,
                               void* stream);
/* Workspace bytes for ap_backproject_matched to take every upright view of
 * a call with this config, view count and sinogram shape in one pass (0 when
 * the gather formulation does not apply: nv * ssv above 2014).             */
uint64_t ap_backproject_matched_workspace_size(const ap_config_cb_project* cfg,
                                               uint32_t num_projs,
                                               const ap_array3* sino_desc);
/* ---- matched adjoint that allocates nothing -----------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ap_backproject_matched without the workspace argument.
 * The gather's prefix sums go to a 64 MiB scratch array that the library
 * holds on each GPU it runs on, from the first call there until the process
 * ends, taken in view chunks that fit it (each chunk adds to vol; results
 * agree with ap_backproject_matched to FP32 rounding).  The call allocates
 * no device memory.  Returns AP_ERR_UNSUPPORTED, with a message, when one
 * view of the call needs more than 64 MiB
 * (ap_backproject_matched_workspace_size(cfg, 1, sino_desc) > 2^26).  All
 * calls on one GPU share its scratch: calls that overlap in time (on other
 * streams or host threads) race.                                           */
int32_t ap_backproject_matched_noalloc(const ap_config_cb_project* cfg,
                                       const ap_proj_spec* specs, uint32_t num_projs,
                                       const ap_vol_tex_chunk_spec* chunk,
                                       const ap_array3* vol_desc, void* vol,
                                       const ap_array3* mult_desc, const void* multiplier,
                                       const ap_array3* sino_desc, const void* sino,
                                       float alpha, void* stream);

/* ---- analytic FORBILD forward projection -------------------------------
 * Replaces the project_CB_FORBILD_shapes launch in cb/flatpanel.py:1042-1109
 * (CB_project_texture_chunked.cu:421-593, compositing
 * cuda/structs/cu/FORBILD_Shape.cu:474-540): last-wins (or additive when
 * overlap_addition) exact line integrals through clipped primitives.
 * sino[p][j][k] += multiply_before_adding * mean_ss integral.
 * Rays meeting more than 1000 shapes add 1 to *out_of_mem_count (device
 * uint32, may be NULL).  Sinogram dtype F32 or F64.                         */
int32_t ap_project_forbild(const ap_config_cb_project_forbild* cfg,
                           const ap_proj_spec* specs, uint32_t num_projs,
                           const ap_forbild_shape* shapes, uint32_t num_shapes,
                           const ap_array3* sino_desc, void* sino,
                           uint32_t* out_of_mem_count, void* stream);

/* ---- FORBILD point-sampling voxeliser ----------------------------------
 * Replaces add_FORBILD_sampling (cuda/cusrc_templates/
 * sample_FORBILD_volumes.cu:10-123): vol[i] += mean over ss^3 sub-points of
 * rho(last shape containing the point).  chunk gives voxwid/texpos and the
 * region written.                                                          */
int32_t ap_voxelise_forbild(const ap_forbild_shape* shapes, uint32_t num_shapes,
                            const ap_vol_tex_chunk_spec* chunk,
                            const ap_array3* vol_desc, void* vol,
                            uint32_t supersample, void* stream);

/* Per-shape axis-aligned bounding boxes (clip planes ignored), replacing
 * get_bboxes_FORBILD_shapes (cuda/cusrc_templates/get_bboxes_FORBILD.cu:8-21).
 * bboxes: device float[6*num_shapes] = (xmin,ymin,zmin,xmax,ymax,zmax).     */
int32_t ap_forbild_bboxes(const ap_forbild_shape* shapes, uint32_t num_shapes,
                          float* bboxes, void* stream);

/* ---- row-wise ramp filter ----------------------------------------------
 * Replaces RampFilterer._filter_chunk_axis1 (cb/ramp_filter.py:139-240) and,
 * with fdk_cosine_weight != 0, the FDK pre-weight of FBPFilterer
 * (ramp_filter.py:407-436).  Linear ("same") convolution along axis 1 (u):
 *   out[p][j][k] = sum_m w(m,k) in[p][m][k] tap(j-m)
 * with tap from `kernel` for pixel width `physpixwid`, and
 * w(m,k) = 1/sqrt(1 + ((m-(nu-1)/2) pw)^2 + ((k-(nv-1)/2) pw)^2) when the
 * cosine weight is on (pw = physpixwid), else 1.  in and out must not
 * overlap; both F32.                                                        */
int32_t ap_ramp_filter_rows(const ap_array3* in_desc, const void* in,
                            const ap_array3* out_desc, void* out,
                            double physpixwid, int32_t kernel,
                            int32_t fdk_cosine_weight, void* stream);

/* ---- FBP detector rotation (GBC sampling) -------------------------------
 * The per-projection steps of FBPFilterer.filter_bulk_axis1_stream
 * (cb/ramp_filter.py:577-886) around the row filter, for in [V][nu][nv]
 * and the expanded buffer [V][Eu][Ev] (Eu - nu and Ev - nv even, the
 * padding split equally on both sides):
 *
 * ap_fbp_rotate_weight: out[p] = W_p * rot(-zeta_p) pad(in[p])  (:612-754)
 *   pad: zero padding to Eu x Ev; rot(a): rotation by angle a (radians)
 *   about the buffer centre ((Eu-1)/2, (Ev-1)/2), i.e. out(o) = in(M(o - c)
 *   + c) with M = [[cos a, sin a], [-sin a, cos a]], sampled bilinearly
 *   with zeros outside (scipy.ndimage.rotate order=1, mode grid-constant,
 *   reshape=False); zeta_p = angles[p] (device float[V]; NULL = all 0;
 *   zeta 0 copies exactly).  W_p(u index i, v index k): with
 *   u = pw (i - (Eu-1)/2), v = pw (k - (Ev-1)/2), pw = pix_over_L,
 *     AP_FBP_GBC, AP_FBP_FAN: 1 / ((1 + v^2)^(3/2) sqrt(1 + u^2 + v^2))
 *     AP_FBP_FDK:             1 / sqrt(1 + u^2 + v^2)
 *   and for AP_FBP_GBC times the letterbox of zeta_p (:470-501, 689-754):
 *   the v range [m - d, m + d], m = (Ev-1)/2,
 *   d = sqrt(max(cos^2 zeta_p / cos^2 zeta_max - 1, 0)) / pw, columns
 *   outside round(m - d) .. round(m + d) zeroed and the end ones weighted by their
 *   covered fraction.
 * ap_fbp_unrotate: out[p] = crop(rot(+zeta_p) in[p])  (:807-886), the
 *   inverse rotation of the filtered expanded buffer, cropped back to
 *   nu x nv.
 * in and out are F32 and must not overlap.                                 */
int32_t ap_fbp_rotate_weight(const ap_array3* in_desc, const void* in,
                             const ap_array3* out_desc, void* out,
                             const float* angles, double pix_over_L,
                             double zeta_max, int32_t mode, void* stream);
int32_t ap_fbp_unrotate(const ap_array3* in_desc, const void* in,
                        const ap_array3* out_desc, void* out,
                        const float* angles, void* stream);

/* ---- GBC weight integral -------------------------------------------------
 * Replaces get_GBC_weight_integ (cuda/cusrc_templates/CB_backproject.cu:
 * 1111-1168, host gbc/weightinteg.py:24-59): for each r = r_on_R[i]
 * (device float[n], increasing), adds to out[i] (device double[n])
 *   4 pi zmax / (nz nt) * sum_{a<nt} sum_{b<nz} h_b w(a, b)
 * with w the preset-7 weight (R = 1, soft_vartheta_* of cfg) of the point
 * (r, 0, 0) seen from the source (cos t_a, sin t_a, z_b), t_a = pi a / nt,
 * z_b = zmax b / (nz - 1), zmax = tan(soft_vartheta_v) (1 + r_on_R[n-1]),
 * h_0 = 1/2 and h_b = 1 otherwise.  nz >= 2.                              */
int32_t ap_gbc_weight_integ(const float* r_on_R, uint32_t n, double* out,
                            uint32_t num_z_samps, uint32_t num_ang_samps,
                            const ap_config_cb_backproject* cfg, void* stream);

/* ---- element-wise volume / sinogram operations --------------------------
 * Replace the texture kernels of cuda/cusrc_templates/texture_operations.cu
 * (fill :124-135, softmask :137-162, invert :164-185, multiply :187-199,
 * inner product :284-307, multiply/add surfaces :309-369).                 */
int32_t ap_array_fill(const ap_array3* desc, void* a, double value, void* stream);
int32_t ap_array_scale(const ap_array3* desc, void* a, double factor, void* stream);
/* a = alpha*a + beta*b  (b may have another dtype, same shape) */
int32_t ap_array_axpby(const ap_array3* a_desc, void* a, double alpha,
                       const ap_array3* b_desc, const void* b, double beta,
                       void* stream);
/* a = a * b element-wise */
int32_t ap_array_mul(const ap_array3* a_desc, void* a,
                     const ap_array3* b_desc, const void* b, void* stream);
/* a = 1/(a + eps*exp(-a/eps)) (eps>0) or 1/a; zeros kept if ignore_zeros */
int32_t ap_array_invert(const ap_array3* desc, void* a, int32_t ignore_zeros,
                        float eps, void* stream);
/* raised-cosine soft mask to the x/y-inscribed cylinder from relrad_start */
int32_t ap_volume_softmask_cylinder(const ap_array3* desc, void* a,
                                    float relrad_start, void* stream);
/* *out (device double) = sum a*b, accumulated in FP64 (a, b same shape).
 * The call overwrites *out. */
int32_t ap_array_dot(const ap_array3* a_desc, const void* a,
                     const ap_array3* b_desc, const void* b,
                     double* out, void* stream);
/* grad += weight * d/dx sum_i sqrt(|grad_fwd x_i|^2 + eps^2)  (isotropic TV,
 * forward differences, Neumann boundary).  x: F32, F16 or BF16 volume;
 * grad: F32.  No reference counterpart (aprict has no TV term). */
int32_t ap_tv_gradient(const ap_array3* x_desc, const void* x,
                       const ap_array3* g_desc, void* grad,
                       float weight, float eps, void* stream);
/* *out (device double) = sum_i sqrt(|grad_fwd x_i|^2 + eps^2), the value
 * whose gradient ap_tv_gradient adds, accumulated in FP64 (overwrites *out). */
int32_t ap_tv_value(const ap_array3* x_desc, const void* x, float eps,
                    double* out, void* stream);

/* ---- fused SIRT iteration ----------------------------------------------
 * One iterate() of iterators/sirt_vram.py:130-149 with the residual fused
 * into the projector epilogue:  r = b - A x  (one write of r, no copy), then
 * x += M (.) BP_cfg(r).  r is caller-owned scratch with b's shape and dtype;
 * b and r are each addressed through their own descriptor (any strides).    */
int32_t ap_sirt_iterate(const ap_config_cb_backproject* bp_cfg,
                        const ap_proj_spec* specs, uint32_t num_projs,
                        const ap_vol_tex_chunk_spec* chunk,
                        const ap_array3* vol_desc, void* x,
                        const ap_array3* m_desc, const void* M,
                        const ap_array3* b_desc, const void* b,
                        const ap_array3* r_desc, void* r,
                        void* stream);

/* ---- launch-shape tuning -------------------------------------------------
 * No counterpart in the reference (its launch shapes are fixed).  For the
 * kernels whose best launch shape depends on the problem size, the library
 * holds one variant index per operation and per host thread (a call reads
 * the calling thread's selection); every variant computes bit-identical
 * results, only the work decomposition differs.  The project-written
 * autotuner (aprict_amd/autotune.py) times the variants per problem class
 * and selects one before launching (also while a HIP graph is being
 * captured: the variant is read when the launch is recorded).
 *   AP_TUNE_PROJECT_RAY     per-ray forward (tilted detectors): pixels per
 *                           workgroup 64x4, 64x2, 64x1 (gathers per batch
 *                           4, 4, 8)
 *   AP_TUNE_PROJECT_COLUMN  column forward (upright detectors): detector rows
 *                           per workgroup 128 (default), 256, 64
 *   AP_TUNE_BACKPROJECT     voxel-driven backprojector: voxels per workgroup
 *                           64x4, 64x2, 64x1                                 */
#define AP_TUNE_PROJECT_RAY     0
#define AP_TUNE_PROJECT_COLUMN  1
#define AP_TUNE_BACKPROJECT     2
#define AP_TUNE_NUM_OPS         3
/* number of variants of op (or -1 for an unknown op) */
int32_t ap_tune_count(int32_t op);
/* selected variant of op (or -1) */
int32_t ap_tune_get(int32_t op);
/* select variant 0 <= variant < ap_tune_count(op) */
int32_t ap_tune_set(int32_t op, int32_t variant);

#ifdef __cplusplus
}
#endif

#endif /* APRICT_HIP_H */
