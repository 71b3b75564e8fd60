// abi.hip -- error reporting, struct-layout checks and descriptor validation
// of the C ABI declared in include/aprict_hip.h.
#include <stdarg.h>
#include <stddef.h>
#include <stdio.h>

#include "ap_common.h"

// Layouts of the reference's kernel-argument structs (natural alignment,
// LP64), fixed at compile time instead of being discovered at run time.
static_assert(sizeof(ap_det_spec) == 48, "det_spec layout");
static_assert(sizeof(ap_proj_spec) == 60, "proj_spec layout");
static_assert(offsetof(ap_proj_spec, dspec) == 12, "proj_spec.dspec offset");
static_assert(sizeof(ap_vol_tex_chunk_spec) == 40, "vol_tex_chunk_spec layout");
static_assert(sizeof(ap_config_cb_project) == 16, "config_CB_project layout");
static_assert(sizeof(ap_config_cb_project_forbild) == 16, "config_CB_project_FORBILD layout");
static_assert(sizeof(ap_config_cb_backproject) == 60, "config_CB_backproject layout");
static_assert(offsetof(ap_config_cb_backproject, mask_to_cyl_rad) == 4, "cfg_bp.mask_to_cyl_rad");
static_assert(offsetof(ap_config_cb_backproject, soften_edge_proportion) == 12, "cfg_bp.soften");
static_assert(offsetof(ap_config_cb_backproject, ssx) == 32, "cfg_bp.ssx");
static_assert(sizeof(ap_forbild_clip) == 20, "FORBILD_ClipAxisSpec layout");
static_assert(sizeof(ap_forbild_shape) == 156, "FORBILD_Shape layout");
static_assert(offsetof(ap_forbild_shape, x) == 68, "FORBILD_Shape.x offset");
static_assert(offsetof(ap_forbild_shape, a_x) == 108, "FORBILD_Shape.a_x offset");
static_assert(sizeof(ap_array3) == 40, "ap_array3 layout");
static_assert(sizeof(ap_proj_chunk_spec_maxlen10) == 640, "proj_chunk_spec_maxlen10 layout");
static_assert(offsetof(ap_proj_chunk_spec_maxlen10, numprojs) == 600, "chunk10.numprojs offset");
static_assert(offsetof(ap_proj_chunk_spec_maxlen10, projmemaddr) == 608, "chunk10.projmemaddr offset");
static_assert(offsetof(ap_proj_chunk_spec_maxlen10, projmemstrides) == 616, "chunk10.projmemstrides offset");

static thread_local char g_last_error[512] = "no error";

int32_t ap_fail(int32_t code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

int32_t ap_check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return ap_fail(AP_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    return AP_OK;
}

int32_t ap_check_array3(const ap_array3* d, const void* ptr, const char* name,
                        bool allow_f16, bool allow_bf16, bool allow_f64) {
    if (!d) return ap_fail(AP_ERR_INVALID, "%s: NULL descriptor", name);
    // an array with no elements may have no storage (as empty torch tensors do)
    if (!ptr && array3_numel(d) != 0) return ap_fail(AP_ERR_INVALID, "%s: NULL data pointer", name);
    switch (d->dtype) {
    case AP_F32: break;
    case AP_F16: if (!allow_f16) return ap_fail(AP_ERR_UNSUPPORTED, "%s: float16 not supported here", name); break;
    case AP_BF16: if (!allow_bf16) return ap_fail(AP_ERR_UNSUPPORTED, "%s: bfloat16 not supported here", name); break;
    case AP_F64: if (!allow_f64) return ap_fail(AP_ERR_UNSUPPORTED, "%s: float64 not supported here", name); break;
    default: return ap_fail(AP_ERR_INVALID, "%s: unknown dtype %d", name, d->dtype);
    }
    for (int a = 0; a < 3; a++)
        if (d->strides[a] < 0)
            return ap_fail(AP_ERR_INVALID, "%s: negative stride", name);
    return AP_OK;
}

namespace {
constexpr int32_t kTuneCount[AP_TUNE_NUM_OPS] = {3, 3, 3};
// one selection per host thread: threads that tune or capture different
// problems at the same time do not see each other's variants
thread_local int32_t g_tune[AP_TUNE_NUM_OPS] = {};
}  // namespace

int32_t ap_tune_variant(int32_t op) { return (op >= 0 && op < AP_TUNE_NUM_OPS) ? g_tune[op] : 0; }

extern "C" {

int32_t ap_tune_count(int32_t op) { return (op >= 0 && op < AP_TUNE_NUM_OPS) ? kTuneCount[op] : -1; }

int32_t ap_tune_get(int32_t op) { return (op >= 0 && op < AP_TUNE_NUM_OPS) ? ap_tune_variant(op) : -1; }

int32_t ap_tune_set(int32_t op, int32_t variant) {
    if (op < 0 || op >= AP_TUNE_NUM_OPS) return ap_fail(AP_ERR_INVALID, "unknown tuning op %d", op);
    if (variant < 0 || variant >= kTuneCount[op])
        return ap_fail(AP_ERR_INVALID, "tuning op %d has %d variants, got %d", op, kTuneCount[op], variant);
    g_tune[op] = variant;
    return AP_OK;
}

const char* ap_last_error(void) { return g_last_error; }

int32_t ap_version(void) { return 100; }

int32_t ap_struct_size(int32_t which) {
    switch (which) {
    case 0: return (int32_t)sizeof(ap_det_spec);
    case 1: return (int32_t)sizeof(ap_proj_spec);
    case 2: return (int32_t)sizeof(ap_vol_tex_chunk_spec);
    case 3: return (int32_t)sizeof(ap_config_cb_project);
    case 4: return (int32_t)sizeof(ap_config_cb_project_forbild);
    case 5: return (int32_t)sizeof(ap_config_cb_backproject);
    case 6: return (int32_t)sizeof(ap_forbild_clip);
    case 7: return (int32_t)sizeof(ap_forbild_shape);
    case 8: return (int32_t)sizeof(ap_array3);
    case 9: return (int32_t)sizeof(ap_proj_chunk_spec_maxlen10);
    default: return -1;
    }
}

}  // extern "C"
