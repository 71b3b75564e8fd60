"""Row-wise ramp filtering (drop-in for ``aprict.cb.ramp_filter``).

``RampFilterer`` mirrors ``src/aprict/cb/ramp_filter.py:36-310``: 1-D filtering
of projection arrays along axis 1 (u) with Ram-Lak (default) or sinc-corrected
ramp taps of pixel width ``physpixwid``.  The reference multiplies FFT spectra
in FP32 batches of 2-3 projections; this version evaluates the identical
"same"-mode linear convolution directly in one HIP launch over all
projections (``ap_ramp_filter_rows``).

``FBPFilterer`` mirrors ``ramp_filter.py:313-888`` for the FDK path
(``_hacky_fdk_mode=True``): cosine pre-weight ``1/sqrt(1+u^2+v^2)`` on the
detector scaled to unit source distance, fused into the same kernel, then
the Ram-Lak filter with pixel width ``physpixwid/L``.  The GBC branch of the
reference (random in-plane detector rotation, letterbox weighting, Hilbert
output) is not part of this library.
"""

from __future__ import annotations

import typing

import numpy as np

from .. import _abi, ramp_kernels

__all__ = ["FBPFilterer", "RampFilterer", "fdk_reconstruct"]


def _torch():
    import torch  # noqa: PLC0415

    return torch


def _kernel_code(fn) -> int:
    if fn is ramp_kernels.ramp_kernel_ramlak:
        return _abi.AP_RAMP_RAMLAK
    if fn is ramp_kernels.ramp_kernel:
        return _abi.AP_RAMP_SINC
    msg = f"kernel function {getattr(fn, '__name__', fn)} has no GPU implementation"
    raise NotImplementedError(msg)


def ramp_filter_rows(arr, physpixwid: float, *, kernel: int = _abi.AP_RAMP_RAMLAK,
                     fdk_cosine_weight: bool = False, out=None):
    """Filter a ``[V, nu, nv]`` float32 GPU tensor along axis 1; returns ``out``."""
    torch = _torch()
    _abi.require_device(arr, "projection array")
    a = arr if arr.dtype == torch.float32 else arr.float()
    if out is None:
        out = torch.empty(a.shape, dtype=torch.float32, device=a.device)
    _abi.check(_abi.lib().ap_ramp_filter_rows(_abi.array3(a), _abi.ptr(a), _abi.array3(out), _abi.ptr(out),
                                              float(physpixwid), int(kernel), int(bool(fdk_cosine_weight)),
                                              _abi.stream_handle(a.device)), "ramp_filter_rows")
    return out


class RampFilterer:
    """Repeated 1-D ramp filtering along axis 1 of ``[V, nu, nv]`` projection arrays."""

    def __init__(self, *, physpixwid: float = 1, proj_2d_shape: tuple[int, int],
                 projs_per_batch: int = 2, stream=None,
                 override_kern_funcs: tuple[typing.Callable, ...] | None = None) -> None:
        self.chunk_shape = (projs_per_batch, *proj_2d_shape)
        self.physpixwid = physpixwid
        self.stream = stream
        funcs = override_kern_funcs or (ramp_kernels.ramp_kernel_ramlak,)
        self._codes = [_kernel_code(f) for f in funcs]

    @property
    def num_kerns(self) -> int:
        return len(self._codes)

    def _check(self, arr) -> None:
        if tuple(arr.shape[1:]) != tuple(self.chunk_shape[1:]):
            msg = "Input projection array shape does not match the shape initialised for this filtering object."
            raise ValueError(msg)

    def filter_bulk_axis1_concat(self, cparr_view, block: bool = True):
        """([one filtered array per kernel], completion event)."""
        torch = _torch()
        if not isinstance(cparr_view, torch.Tensor):
            cparr_view = torch.from_numpy(np.ascontiguousarray(cparr_view, dtype=np.float32)).cuda()
        self._check(cparr_view)
        outs = [ramp_filter_rows(cparr_view, self.physpixwid, kernel=c) for c in self._codes]
        ev = torch.cuda.Event()
        ev.record()
        if block:
            ev.synchronize()
        return outs, ev

    def filter_bulk_axis1_stream(self, cparr_view, block: bool = True):
        """Generator over batches of ``projs_per_batch`` projections."""
        n = self.chunk_shape[0]
        for i in range(0, cparr_view.shape[0], n):
            yield self.filter_bulk_axis1_concat(cparr_view[i:i + n], block=block)


class FBPFilterer:
    """FDK filtering: cosine weight + Ram-Lak along u at pixel width ``physpixwid / L``."""

    def __init__(self, *, L: float, physpixwid: float, proj_2d_shape, projs_per_batch: int = 3,  # noqa: N803
                 sampling_strategy=None, limit_zeta: float | None = None,
                 include_hilbert_transform_in_output: bool = False, _hacky_fan_beam_mode: bool = False,
                 _hacky_fdk_mode: bool = False, rng=None) -> None:
        del sampling_strategy, limit_zeta, rng
        if not _hacky_fdk_mode or _hacky_fan_beam_mode or include_hilbert_transform_in_output:
            msg = "only the FDK filtering mode (_hacky_fdk_mode=True, no Hilbert output) is implemented"
            raise NotImplementedError(msg)
        self.L = L
        self.physpixwid = physpixwid
        self.proj_2d_shape = (int(proj_2d_shape[0]), int(proj_2d_shape[1]))
        self.chunk_shape = (projs_per_batch, *self.proj_2d_shape)
        self.hacky_fdk_mode = True

    def filter_bulk_axis1_concat(self, arr, block: bool = True, force_rotation_angle: float | None = None):
        torch = _torch()
        if force_rotation_angle not in (None, 0, 0.0):
            msg = "detector rotation (GBC sampling) is not implemented"
            raise NotImplementedError(msg)
        if not isinstance(arr, torch.Tensor):
            arr = torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float32)).cuda()
        if tuple(arr.shape[1:]) != self.proj_2d_shape:
            msg = "Wrong input size."
            raise ValueError(msg)
        out = ramp_filter_rows(arr, self.physpixwid / self.L, kernel=_abi.AP_RAMP_RAMLAK,
                               fdk_cosine_weight=True)
        ev = torch.cuda.Event()
        ev.record()
        if block:
            ev.synchronize()
        return [out], ev


def fdk_reconstruct(projs, volume, *, R: float, L: float, physpixwid: float,  # noqa: N803
                    bilinear: bool = True, cyl_mask_rad: float | None = None):
    """FDK as in ``scripts/recon_cbfp_cyl_fbp.py`` (fdk geometry): filter,
    fast direct-raycast backprojection with weight (Z/R)^-2, then x 1/(4 pi R V).

    ``projs`` is a ProjectionSet with data; ``volume`` receives the result
    (added into it).
    """
    from . import flatpanel as cbfp  # noqa: PLC0415

    filt = FBPFilterer(L=L, physpixwid=physpixwid, proj_2d_shape=projs.data.shape[-2:],
                       _hacky_fdk_mode=True)
    (filtered,), _ = filt.filter_bulk_axis1_concat(projs._flat_data(), block=False)
    fp = cbfp.ProjectionSet(projs.proj_geom, data=filtered.view(projs.data.shape))
    cfg = cbfp.ConfigBackproject(fast_approx=True, cyl_mask_rad=cyl_mask_rad, Z_on_val_factor_power=-2,
                                 Z_on_val_factor_val=R, approx_mode_direct_raycast=True,
                                 approx_mode_bilinear_interpolate=bilinear)
    tmp = volume.zeroed_copy()
    cbfp.backproject(cfg, tmp, fp)
    tmp.multiply(1.0 / (4.0 * np.pi * R * projs.proj_geom.num_projs))
    volume.add(tmp)
    return volume
