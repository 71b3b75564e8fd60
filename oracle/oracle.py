"""ctypes wrapper of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker / the timed CPU port; the
product package never imports it.

``liboracle64.so`` is the FP64 oracle; ``liboracle32.so`` the FP32 OpenMP port
of the reference semantics (same functions with an ``_f32`` suffix).
Arrays are dense C-order numpy arrays: volumes [Nx][Ny][Nz], sinograms
[V][nu][nv]; descriptors are numpy structured scalars with the ABI dtypes.
"""

from __future__ import annotations

import ctypes
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
# ORACLE_BUILD selects another build of the same sources (e.g. _build_san,
# the AddressSanitizer/UBSan build of `make sanitize`)
BUILD = HERE / os.environ.get("ORACLE_BUILD", "_build")

_P = ctypes.c_void_p
_LIBS: dict[bool, ctypes.CDLL] = {}


def build() -> None:
    """Compile both oracle libraries (make in oracle/)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def use_native_port(outdir: str) -> str:
    """Build the FP32 port with -march=native for this host into ``outdir``
    and use it for the fp32 functions from now on.  Returns what was built
    ("native") or why the portable x86-64-v3 build stays in use."""
    try:
        subprocess.run(["make", "-s", "-C", str(HERE), "native", f"NATIVE_OUT={outdir}"], check=True,
                       capture_output=True, text=True, timeout=300)
        _LIBS.pop(True, None)
        _load(True, pathlib.Path(outdir) / "liboracle32_native.so")
        return "native"
    except (OSError, subprocess.SubprocessError) as e:
        return f"x86-64-v3 (native build failed: {e})"


def _lib(fp32: bool) -> ctypes.CDLL:
    if fp32 not in _LIBS:
        path = BUILD / ("liboracle32.so" if fp32 else "liboracle64.so")
        if not path.exists():
            build()
        _load(fp32, path)
    return _LIBS[fp32]


def _load(fp32: bool, path: pathlib.Path) -> None:
    lib = ctypes.CDLL(str(path))
    real = ctypes.c_float if fp32 else ctypes.c_double
    sfx = "_f32" if fp32 else ""
    getattr(lib, "or_slab_length" + sfx).restype = real
    getattr(lib, "or_gbc_weight" + sfx).restype = real
    for name in ("or_project_voxel", "or_backproject", "or_backproject_matched", "or_ramp_filter",
                 "or_project_forbild", "or_voxelise_forbild", "or_project_voxel_refdda", "or_backproject_gbc",
                 "or_gbc_weight_integ"):
        getattr(lib, name + sfx).restype = ctypes.c_int
    _LIBS[fp32] = lib


def _fn(name: str, fp32: bool):
    return getattr(_lib(fp32), name + ("_f32" if fp32 else ""))


def _real(fp32: bool):
    return np.float32 if fp32 else np.float64


def _p(a) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)


def _dense(a, fp32: bool) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=_real(fp32))


def _dims(vol: np.ndarray) -> np.ndarray:
    return np.array(vol.shape, dtype=np.uint32)


def _ok(st: int, what: str) -> None:
    if st != 0:
        msg = f"oracle {what} failed with status {st}"
        raise RuntimeError(msg)


def project_voxel(cfg, specs, chunk, vol, sino, b=None, *, nthreads: int = 0, fp32: bool = False):
    """sino (+)= mba * mean_ss integral; returns the (new) sinogram array."""
    vol = _dense(vol, fp32)
    sino = _dense(sino, fp32).copy()
    bb = _dense(b, fp32) if b is not None else None
    specs = np.ascontiguousarray(specs).reshape(-1)
    st = _fn("or_project_voxel", fp32)(_p(cfg), _p(specs), ctypes.c_uint32(len(specs)), _p(chunk),
                                       _p(_dims(vol)), _p(vol), _p(sino), _p(bb), ctypes.c_int(nthreads))
    _ok(st, "project_voxel")
    return sino


def project_voxel_refdda(cfg, specs, chunk, vol, sino, *, nthreads: int = 0):
    """FP32 port of the reference traversal (adds into a copy of sino)."""
    vol = _dense(vol, True)
    sino = _dense(sino, True).copy()
    specs = np.ascontiguousarray(specs).reshape(-1)
    st = _fn("or_project_voxel_refdda", True)(_p(cfg), _p(specs), ctypes.c_uint32(len(specs)), _p(chunk),
                                              _p(_dims(vol)), _p(vol), _p(sino), ctypes.c_int(nthreads))
    _ok(st, "project_voxel_refdda")
    return sino


def backproject(cfg, specs, chunk, vol, sino, mult=None, *, nthreads: int = 0, fp32: bool = False):
    """vol + mult * BP(sino) with the reference's backprojector semantics."""
    vol = _dense(vol, fp32).copy()
    sino = _dense(sino, fp32)
    m = _dense(mult, fp32) if mult is not None else None
    specs = np.ascontiguousarray(specs).reshape(-1)
    st = _fn("or_backproject", fp32)(_p(cfg), _p(specs), ctypes.c_uint32(len(specs)), _p(chunk),
                                     _p(_dims(vol)), _p(vol), _p(m), _p(sino), ctypes.c_int(nthreads))
    _ok(st, "backproject")
    return vol


def backproject_gbc(cfg, specs, chunk, vol, sec, sino, mult=None, *, nthreads: int = 0, fp32: bool = False):
    """(vol + mult * sum_views w y, sec + mult * sum_views w): the preset-7 GBC backprojection."""
    vol = _dense(vol, fp32).copy()
    sec = _dense(sec, fp32).copy()
    sino = _dense(sino, fp32)
    m = _dense(mult, fp32) if mult is not None else None
    specs = np.ascontiguousarray(specs).reshape(-1)
    st = _fn("or_backproject_gbc", fp32)(_p(cfg), _p(specs), ctypes.c_uint32(len(specs)), _p(chunk),
                                         _p(_dims(vol)), _p(vol), _p(sec), _p(m), _p(sino), ctypes.c_int(nthreads))
    _ok(st, "backproject_gbc")
    return vol, sec


def gbc_weight(cfg, S, x) -> float:
    """Preset-7 weight of the source S for the point x (R = Z_on_val_factor_val)."""
    s3 = np.ascontiguousarray(S, dtype=np.float64)
    x3 = np.ascontiguousarray(x, dtype=np.float64)
    return float(_fn("or_gbc_weight", False)(_p(cfg), _p(s3), _p(x3)))


def gbc_weight_integ(r_on_R, cfg, num_z_samps: int, num_ang_samps: int) -> np.ndarray:
    """Quadrature of the preset-7 weight over the source cylinder (R = 1) for points (r, 0, 0)."""
    r = np.ascontiguousarray(r_on_R, dtype=np.float64)
    out = np.zeros_like(r)
    st = _fn("or_gbc_weight_integ", False)(_p(r), ctypes.c_uint32(len(r)), _p(out), ctypes.c_uint32(num_z_samps),
                                           ctypes.c_uint32(num_ang_samps), _p(cfg))
    _ok(st, "gbc_weight_integ")
    return out


def backproject_matched(cfg, specs, chunk, vol, sino, mult=None, alpha: float = 1.0, *, fp32: bool = False):
    """vol + alpha * mult * A^T sino (A = project_voxel with cfg)."""
    vol = _dense(vol, fp32).copy()
    sino = _dense(sino, fp32)
    m = _dense(mult, fp32) if mult is not None else None
    specs = np.ascontiguousarray(specs).reshape(-1)
    real = ctypes.c_float if fp32 else ctypes.c_double
    st = _fn("or_backproject_matched", fp32)(_p(cfg), _p(specs), ctypes.c_uint32(len(specs)), _p(chunk),
                                             _p(_dims(vol)), _p(vol), _p(m), _p(sino), real(alpha))
    _ok(st, "backproject_matched")
    return vol


def slab_length(S, Q, lo, hi) -> float:
    """Length of segment S->Q inside the box [lo, hi] (brute force)."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (S, Q, lo, hi)]
    return float(_fn("or_slab_length", False)(*[_p(a) for a in arrs]))


def ramp_filter(arr, physpixwid: float, *, kernel: int = 0, fdk: bool = False, fp32: bool = False):
    """Direct linear convolution along axis 1 (long double accumulation)."""
    a = _dense(arr, fp32)
    out = np.empty_like(a)
    v, nu, nv = a.shape
    st = _fn("or_ramp_filter", fp32)(ctypes.c_uint32(v), ctypes.c_uint32(nu), ctypes.c_uint32(nv), _p(a),
                                     _p(out), ctypes.c_double(physpixwid), ctypes.c_int(kernel),
                                     ctypes.c_int(int(fdk)))
    _ok(st, "ramp_filter")
    return out


def project_forbild(cfg_fb, specs, shapes, sino, *, nthreads: int = 0, fp32: bool = False):
    """sino + mba * mean_ss last-wins analytic integral."""
    sino = _dense(sino, fp32).copy()
    shapes = np.ascontiguousarray(shapes).reshape(-1)
    specs = np.ascontiguousarray(specs).reshape(-1)
    st = _fn("or_project_forbild", fp32)(_p(cfg_fb), _p(specs), ctypes.c_uint32(len(specs)), _p(shapes),
                                         ctypes.c_uint32(len(shapes)), _p(sino), ctypes.c_int(nthreads))
    _ok(st, "project_forbild")
    return sino


def voxelise_forbild(shapes, chunk, vol, ss: int = 1, *, nthreads: int = 0, fp32: bool = False):
    vol = _dense(vol, fp32).copy()
    shapes = np.ascontiguousarray(shapes).reshape(-1)
    st = _fn("or_voxelise_forbild", fp32)(_p(shapes), ctypes.c_uint32(len(shapes)), _p(chunk),
                                          _p(_dims(vol)), _p(vol), ctypes.c_uint32(ss), ctypes.c_int(nthreads))
    _ok(st, "voxelise_forbild")
    return vol


# ---------------------------------------------------------------------------
# smoothed isotropic total variation (numpy, FP64).  aprict has no TV term:
# the definition is the one the differentiable workload states,
# TV_eps(x) = sum_i sqrt(|D x_i|^2 + eps^2) with forward differences D that
# are zero across the far faces.  Pinned in tests/test_oracle_tv.py by
# closed forms and by central differences of tv_value.
# ---------------------------------------------------------------------------
def _tv_diffs(x: np.ndarray):
    x = np.asarray(x, dtype=np.float64)
    dx = np.zeros_like(x)
    dy = np.zeros_like(x)
    dz = np.zeros_like(x)
    dx[:-1] = x[1:] - x[:-1]
    dy[:, :-1] = x[:, 1:] - x[:, :-1]
    dz[:, :, :-1] = x[:, :, 1:] - x[:, :, :-1]
    return dx, dy, dz


def tv_value(x: np.ndarray, eps: float) -> float:
    dx, dy, dz = _tv_diffs(x)
    return float(np.sqrt(dx * dx + dy * dy + dz * dz + eps * eps).sum())


def tv_gradient(x: np.ndarray, eps: float) -> np.ndarray:
    """d tv_value / d x (the adjoint of D applied to D x / |D x|_eps)."""
    dx, dy, dz = _tv_diffs(x)
    n = np.sqrt(dx * dx + dy * dy + dz * dz + eps * eps)
    px, py, pz = dx / n, dy / n, dz / n
    g = -(px + py + pz)
    g[1:] += px[:-1]
    g[:, 1:] += py[:, :-1]
    g[:, :, 1:] += pz[:, :, :-1]
    return g
