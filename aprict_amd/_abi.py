"""ctypes mirror of ``include/aprict_hip.h`` and the loader of libaprict_hip.so.

The POD structs below reproduce, field for field, the kernel-argument structs
of aprict (``src/aprict/cuda/structs/geometry.py:40-186``,
``config_cb_project.py:27-39``, ``config_cb_project_FORBILD.py:28-40``,
``config_cb_backproject.py:27-72``, ``FORBILD_volumes.py:502-607``).  Each has
a numpy structured dtype with the same layout, so arrays built for the
reference can be handed to this library unchanged.  Layouts are checked
against the library's own ``ap_struct_size`` at load time.
"""

from __future__ import annotations

import ctypes
import pathlib
import threading

import numpy as np

_LIB_PATH = pathlib.Path(__file__).resolve().parent / "lib" / "libaprict_hip.so"

AP_OK = 0
AP_ERR_INVALID = -1
AP_ERR_UNSUPPORTED = -2
AP_ERR_HIP = -3

AP_F32, AP_F16, AP_BF16, AP_F64 = 0, 1, 2, 3
AP_RAMP_RAMLAK, AP_RAMP_SINC, AP_RAMP_HILBERT, AP_RAMP_HILBERT_HALF = 0, 1, 2, 3
AP_FBP_GBC, AP_FBP_FAN, AP_FBP_FDK = 0, 1, 2


# --------------------------------------------------------------------------
# numpy structured dtypes (natural alignment = the C layout)
# --------------------------------------------------------------------------
fvec3 = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4")], align=True)
uivec3 = np.dtype([("x", "<u4"), ("y", "<u4"), ("z", "<u4")], align=True)
ulvec3 = np.dtype([("x", "<u8"), ("y", "<u8"), ("z", "<u8")], align=True)
det_spec = np.dtype(
    [("pos", fvec3), ("u", fvec3), ("v", fvec3), ("nu", "<u4"), ("nv", "<u4"), ("type", "u1")],
    align=True,
)
proj_spec = np.dtype([("sourcepos", fvec3), ("dspec", det_spec)], align=True)
vol_tex_chunk_spec = np.dtype(
    [("voxwid", "<f4"), ("offsets", uivec3), ("shape", uivec3), ("texpos", fvec3)], align=True
)
config_cb_project = np.dtype(
    [
        ("shrink_volume_factor", "<f4"),
        ("multiply_before_adding", "<f4"),
        ("ssu", "<u4"),
        ("ssv", "<u4"),
    ],
    align=True,
)
config_cb_project_FORBILD = np.dtype(
    [("multiply_before_adding", "<f4"), ("overlap_addition", "?"), ("ssu", "<u4"), ("ssv", "<u4")],
    align=True,
)
config_cb_backproject = np.dtype(
    [
        ("exact_line_integ_backproj", "?"),
        ("backproject_all_1s", "?"),
        ("mask_to_cylinder", "?"),
        ("mask_to_cyl_rad", "<f4"),
        ("apply_summand_multiplier", "?"),
        ("soften_edge_proportion", "<f4"),
        ("Z_on_L_factor_power", "<i4"),
        ("Z_on_val_factor_power", "<i4"),
        ("Z_on_val_factor_val", "<f4"),
        ("approx_mode_direct_raycast", "?"),
        ("approx_mode_linear_interpolate", "?"),
        ("backproj_weighting_preset", "u1"),
        ("ssx", "<u4"),
        ("ssy", "<u4"),
        ("ssz", "<u4"),
        ("soft_vartheta_v", "<f4"),
        ("soft_vartheta_v_softspan", "<f4"),
        ("soft_vartheta_h", "<f4"),
        ("soft_vartheta_h_softspan", "<f4"),
    ],
    align=True,
)
forbild_clip = np.dtype([("axis", fvec3), ("dotprod", "<f4"), ("mode", "u1")], align=True)
forbild_shape = np.dtype(
    [
        ("type", "u1"),
        ("rho", "<f4"),
        ("cas1", forbild_clip),
        ("cas2", forbild_clip),
        ("cas3", forbild_clip),
        ("x", "<f4"),
        ("y", "<f4"),
        ("z", "<f4"),
        ("dx", "<f4"),
        ("dy", "<f4"),
        ("dz", "<f4"),
        ("r", "<f4"),
        ("r1", "<f4"),
        ("r2", "<f4"),
        ("l", "<f4"),
        ("a_x", fvec3),
        ("a_y", fvec3),
        ("a_z", fvec3),
        ("axis", fvec3),
    ],
    align=True,
)

# the reference's by-value argument of one backprojection launch (10 views)
proj_chunk_spec_maxlen10 = np.dtype(
    [("projspecs", proj_spec, (10,)), ("numprojs", "<u1"), ("projmemaddr", "<u8"), ("projmemstrides", ulvec3)],
    align=True,
)

_EXPECTED_SIZES = {
    0: det_spec.itemsize,
    1: proj_spec.itemsize,
    2: vol_tex_chunk_spec.itemsize,
    3: config_cb_project.itemsize,
    4: config_cb_project_FORBILD.itemsize,
    5: config_cb_backproject.itemsize,
    6: forbild_clip.itemsize,
    7: forbild_shape.itemsize,
    8: 40,
    9: proj_chunk_spec_maxlen10.itemsize,
}
assert _EXPECTED_SIZES[0] == 48 and _EXPECTED_SIZES[1] == 60, "numpy det/proj spec layout"
assert _EXPECTED_SIZES[5] == 60 and _EXPECTED_SIZES[7] == 156, "numpy config/shape layout"


class Array3(ctypes.Structure):
    """``ap_array3``: strided 3-D device array descriptor (element strides)."""

    _fields_ = [
        ("dtype", ctypes.c_int32),
        ("shape", ctypes.c_uint32 * 3),
        ("strides", ctypes.c_int64 * 3),
    ]


# Functions exported by the library: name -> (restype, argtypes)
_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_I32 = ctypes.c_int32
_F32 = ctypes.c_float
_F64 = ctypes.c_double
_A3 = ctypes.POINTER(Array3)

SIGNATURES: dict[str, tuple[object, list[object]]] = {
    "ap_last_error": (ctypes.c_char_p, []),
    "ap_version": (_I32, []),
    "ap_struct_size": (_I32, [_I32]),
    "ap_project_voxel": (_I32, [_P, _P, _U32, _P, _A3, _P, _A3, _P, _A3, _P, _P]),
    "ap_backproject": (_I32, [_P, _P, _U32, _P, _A3, _P, _A3, _P, _A3, _P, _A3, _P, _P]),
    "ap_backproject_matched": (_I32, [_P, _P, _U32, _P, _A3, _P, _A3, _P, _A3, _P, _F32, _P, ctypes.c_uint64, _P]),
    "ap_backproject_matched_workspace_size": (ctypes.c_uint64, [_P, _U32, _A3]),
    "ap_backproject_matched_noalloc": (_I32, [_P, _P, _U32, _P, _A3, _P, _A3, _P, _A3, _P, _F32, _P]),
    "ap_project_forbild": (_I32, [_P, _P, _U32, _P, _U32, _A3, _P, _P, _P]),
    "ap_voxelise_forbild": (_I32, [_P, _U32, _P, _A3, _P, _U32, _P]),
    "ap_forbild_bboxes": (_I32, [_P, _U32, _P, _P]),
    "ap_ramp_filter_rows": (_I32, [_A3, _P, _A3, _P, _F64, _I32, _I32, _P]),
    "ap_fbp_rotate_weight": (_I32, [_A3, _P, _A3, _P, _P, _F64, _F64, _I32, _P]),
    "ap_fbp_unrotate": (_I32, [_A3, _P, _A3, _P, _P, _P]),
    "ap_gbc_weight_integ": (_I32, [_P, _U32, _P, _U32, _U32, _P, _P]),
    "ap_array_fill": (_I32, [_A3, _P, _F64, _P]),
    "ap_array_scale": (_I32, [_A3, _P, _F64, _P]),
    "ap_array_axpby": (_I32, [_A3, _P, _F64, _A3, _P, _F64, _P]),
    "ap_array_mul": (_I32, [_A3, _P, _A3, _P, _P]),
    "ap_array_invert": (_I32, [_A3, _P, _I32, _F32, _P]),
    "ap_volume_softmask_cylinder": (_I32, [_A3, _P, _F32, _P]),
    "ap_array_dot": (_I32, [_A3, _P, _A3, _P, _P, _P]),
    "ap_tv_gradient": (_I32, [_A3, _P, _A3, _P, _F32, _F32, _P]),
    "ap_tv_value": (_I32, [_A3, _P, _F32, _P, _P]),
    "ap_sirt_iterate": (_I32, [_P, _P, _U32, _P, _A3, _P, _A3, _P, _A3, _P, _A3, _P, _P]),
    "ap_tune_count": (_I32, [_I32]),
    "ap_tune_get": (_I32, [_I32]),
    "ap_tune_set": (_I32, [_I32, _I32]),
}

# launch-shape tuning ops (include/aprict_hip.h AP_TUNE_*)
TUNE_PROJECT_RAY = 0
TUNE_PROJECT_COLUMN = 1
TUNE_BACKPROJECT = 2
TUNE_NUM_OPS = 3


class AprictError(RuntimeError):
    """Raised when a library call returns a non-zero status."""


class _Lib:
    _lock = threading.Lock()
    _handle: ctypes.CDLL | None = None

    @classmethod
    def get(cls) -> ctypes.CDLL:
        with cls._lock:
            if cls._handle is None:
                if not _LIB_PATH.exists():
                    msg = (
                        f"{_LIB_PATH} is missing: the HIP kernels are not built "
                        "(run `python -m aprict_amd.build`). There is no CPU fallback."
                    )
                    raise AprictError(msg)
                h = ctypes.CDLL(str(_LIB_PATH))
                for name, (res, args) in SIGNATURES.items():
                    fn = getattr(h, name)
                    fn.restype = res
                    fn.argtypes = args
                for which, size in _EXPECTED_SIZES.items():
                    got = h.ap_struct_size(which)
                    if got != size:
                        msg = f"ABI layout mismatch for struct id {which}: library {got} B, binding {size} B"
                        raise AprictError(msg)
                cls._handle = h
            return cls._handle


def lib() -> ctypes.CDLL:
    """The loaded libaprict_hip.so (raises AprictError if it is not built)."""
    return _Lib.get()


def library_path() -> pathlib.Path:
    return _LIB_PATH


def check(status: int, what: str = "") -> None:
    """Raise AprictError (ValueError / NotImplementedError flavoured) on failure."""
    if status == AP_OK:
        return
    msg = lib().ap_last_error().decode(errors="replace")
    full = f"{what}: {msg}" if what else msg
    if status == AP_ERR_INVALID:
        raise ValueError(full)
    if status == AP_ERR_UNSUPPORTED:
        raise NotImplementedError(full)
    raise AprictError(f"{full} (status {status})")


# --------------------------------------------------------------------------
# tensors -> descriptors
# --------------------------------------------------------------------------
def _torch():
    import torch  # noqa: PLC0415  (torch is heavy; imported on first use)

    return torch


def dtype_code(dt) -> int:
    torch = _torch()
    table = {
        torch.float32: AP_F32,
        torch.float16: AP_F16,
        torch.bfloat16: AP_BF16,
        torch.float64: AP_F64,
    }
    if dt not in table:
        msg = f"unsupported tensor dtype {dt}"
        raise NotImplementedError(msg)
    return table[dt]


def array3(t) -> Array3:
    """Descriptor of a 3-D torch tensor (device memory is not checked here)."""
    if t.dim() != 3:
        msg = f"expected a 3-D tensor, got shape {tuple(t.shape)}"
        raise ValueError(msg)
    d = Array3()
    d.dtype = dtype_code(t.dtype)
    for a in range(3):
        d.shape[a] = int(t.shape[a])
        d.strides[a] = int(t.stride(a))
    return d


def require_device(t, what: str) -> None:
    if not t.is_cuda:
        msg = f"{what} must be a GPU (HIP) tensor; this library has no CPU path"
        raise ValueError(msg)


def ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(int(t.data_ptr())) if t is not None else ctypes.c_void_p(0)


def stream_handle(device=None) -> ctypes.c_void_p:
    torch = _torch()
    s = torch.cuda.current_stream(device)
    return ctypes.c_void_p(int(s.cuda_stream))


def struct_ptr(arr: np.ndarray) -> ctypes.c_void_p:
    """Pointer to a numpy structured scalar/array kept alive by the caller."""
    return ctypes.c_void_p(arr.ctypes.data)


# --------------------------------------------------------------------------
# workspaces
# --------------------------------------------------------------------------
_WORKSPACES: dict[tuple[int, int], object] = {}
_WS_LOCK = threading.Lock()


def workspace(nbytes: int, device=None):
    """Device scratch of at least ``nbytes`` bytes for the current stream of
    ``device`` (a cached uint8 tensor, grown on demand; torch allocations are
    512-byte aligned).  Calls that share it are ordered by their stream."""
    torch = _torch()
    if nbytes <= 0:
        return None
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    idx = dev.index if dev.index is not None else torch.cuda.current_device()
    key = (idx, int(torch.cuda.current_stream(idx).cuda_stream))
    with _WS_LOCK:
        buf = _WORKSPACES.get(key)
        if buf is None or buf.numel() < nbytes:
            _WORKSPACES.pop(key, None)
            buf = torch.empty(int(nbytes), dtype=torch.uint8, device=torch.device("cuda", idx))
            _WORKSPACES[key] = buf
        return buf


def matched_workspace(cfg, num_projs: int, sino) -> tuple[ctypes.c_void_p, int]:
    """(pointer, bytes) of the workspace that lets ap_backproject_matched take
    every upright view of the call in one pass (NULL, 0 when it needs none)."""
    nbytes = int(lib().ap_backproject_matched_workspace_size(struct_ptr(cfg), int(num_projs), array3(sino)))
    if nbytes == 0:
        return ctypes.c_void_p(0), 0
    buf = workspace(nbytes, sino.device)
    return ptr(buf), nbytes
