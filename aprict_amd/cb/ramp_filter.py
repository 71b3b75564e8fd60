"""Row-wise ramp filtering and FBP filtering (drop-in for ``aprict.cb.ramp_filter``).

``RampFilterer`` mirrors ``src/aprict/cb/ramp_filter.py:36-310``: 1-D filtering
of projection arrays along axis 1 (u) with Ram-Lak (default), sinc-corrected
ramp or Hilbert taps of pixel width ``physpixwid``.  The reference multiplies
FFT spectra in FP32 batches of 2-3 projections; this version evaluates the
identical "same"-mode linear convolution directly in one HIP launch over all
projections (``ap_ramp_filter_rows``).

``FBPFilterer`` mirrors ``ramp_filter.py:313-936``.  Each projection is
zero-padded to the expanded detector, rotated by a random angle -zeta in
[-zeta_max, zeta_max] (drawn from ``rng``, one ``rng.random()`` per
projection) or by ``force_rotation_angle``, weighted by the cone weight
``1/((1+v^2)^(3/2) sqrt(1+u^2+v^2))`` (cosine weight ``1/sqrt(1+u^2+v^2)`` in
FDK mode) on the detector scaled to unit source distance and, in the GBC
mode, by the letterbox of its angle; then it is filtered along u with the
Ram-Lak taps of pixel width ``physpixwid/L`` (and the Hilbert taps when
``include_hilbert_transform_in_output``), rotated back by +zeta and cropped
(``ap_fbp_rotate_weight``, ``ap_ramp_filter_rows``, ``ap_fbp_unrotate``).
"""

from __future__ import annotations

import enum
import typing

import numpy as np

from .. import _abi, ramp_kernels

__all__ = ["FBPFilterer", "RampFilterer", "fdk_reconstruct"]


def _torch():
    import torch  # noqa: PLC0415

    return torch


def _kernel_code(fn) -> int:
    codes = {ramp_kernels.ramp_kernel_ramlak: _abi.AP_RAMP_RAMLAK, ramp_kernels.ramp_kernel: _abi.AP_RAMP_SINC,
             ramp_kernels.hilbert_transform_kernel: _abi.AP_RAMP_HILBERT,
             ramp_kernels.hilbert_transform_kernel_halfoffset: _abi.AP_RAMP_HILBERT_HALF}
    if fn in codes:
        return codes[fn]
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
    """FBP filtering with random in-plane detector rotations (``cb/ramp_filter.py:313-936``)."""

    class SamplingStrategy(enum.Enum):
        """Strategies to sample the rotation angle of each projection."""

        SINGLE_ROTATION_RANDOM_UNIFORM = 1

    def __init__(self, *, L: float, physpixwid: float, proj_2d_shape, projs_per_batch: int = 3,  # noqa: N803
                 sampling_strategy: SamplingStrategy = SamplingStrategy.SINGLE_ROTATION_RANDOM_UNIFORM,
                 limit_zeta: float | None = None, include_hilbert_transform_in_output: bool = False,
                 _hacky_fan_beam_mode: bool = False, _hacky_fdk_mode: bool = False, rng=None) -> None:
        self.rng = np.random.default_rng(seed=None) if rng is None else rng
        self._include_hilbert_transform = include_hilbert_transform_in_output
        self.hacky_fan_beam_mode = _hacky_fan_beam_mode
        self.hacky_fdk_mode = _hacky_fdk_mode
        self.L = L
        self.physpixwid = physpixwid
        self.ss = sampling_strategy
        self._max_abs_rotation_angle = self.get_max_abs_rotation_angle(proj_2d_shape, physpixwid, L)
        if limit_zeta is not None:
            self._max_abs_rotation_angle = min(self._max_abs_rotation_angle, limit_zeta)
        self.proj_2d_shape = (int(proj_2d_shape[0]), int(proj_2d_shape[1]))
        self._expanded_2d_shape = self.get_expanded_proj_2d_shape(
            proj_2d_shape, self._max_abs_rotation_angle, _hacky_fan_beam_mode, _hacky_fdk_mode)
        self.chunk_shape = (projs_per_batch, *self._expanded_2d_shape)
        kern_funcs = (ramp_kernels.ramp_kernel_ramlak,)
        if include_hilbert_transform_in_output:
            kern_funcs += (ramp_kernels.hilbert_transform_kernel,)
        self.rfilt = RampFilterer(physpixwid=physpixwid / L, proj_2d_shape=self.chunk_shape[1:],
                                  projs_per_batch=self.chunk_shape[0], override_kern_funcs=kern_funcs)
        self.event_completed_filter = None
        self.event_completed_ramp = None

    @property
    def _mode(self) -> int:
        if self.hacky_fdk_mode:
            return _abi.AP_FBP_FDK
        return _abi.AP_FBP_FAN if self.hacky_fan_beam_mode else _abi.AP_FBP_GBC

    def advance_rng(self, amount: int) -> None:
        """Advance the internal rng by this many ``random()`` draws."""
        for _ in range(amount):
            self.rng.random()

    def uniform_rng(self, low: float, high: float, len: int) -> np.ndarray:  # noqa: A002
        """``len`` float32 values ``low + (high - low) * rng.random()``, one draw each."""
        out = np.empty((len,), dtype=np.float32)
        for i in range(len):
            out[i] = low + (high - low) * self.rng.random()
        return out

    def get_letterbox_slice(self, rotation_angle: float) -> slice:
        """Fractional v-index range covered by the detector rotated by ``rotation_angle``."""
        if abs(rotation_angle) > self._max_abs_rotation_angle:
            msg = "Invalid rotation angle entered."
            raise ValueError(msg)
        mid = (self._expanded_2d_shape[1] - 1) / 2
        disc = self.L / self.physpixwid * np.sqrt(
            np.maximum(np.square(np.cos(rotation_angle) / np.cos(self._max_abs_rotation_angle)) - 1, 0))
        return slice(mid - disc, mid + disc)

    @property
    def zeta(self) -> float:
        """The largest rotation angle this object accommodates."""
        return self._max_abs_rotation_angle

    @classmethod
    def get_max_abs_rotation_angle(cls, proj_2d_shape, physpixwid: float, L: float) -> float:  # noqa: N803
        """Half the vertical cone angle at the detector's horizontal edge."""
        nu, nv = int(proj_2d_shape[0]), int(proj_2d_shape[1])
        cone_angle_vert = 2 * np.arctan2(nv * physpixwid / 2, np.sqrt(np.square(L) + np.square(nu * physpixwid / 2)))
        return cone_angle_vert / 2

    @classmethod
    def get_expanded_proj_2d_shape(cls, proj_2d_shape, max_abs_rotation_angle: float,
                                   _hacky_fan_beam_mode: bool = False, _hacky_fdk_mode: bool = False) -> tuple:
        """Detector shape widened along u to hold every rotation (same parity as the original)."""
        if _hacky_fan_beam_mode and _hacky_fdk_mode:
            msg = "Cannot use both _hacky_fan_beam_mode and _hacky_fdk_mode."
            raise ValueError(msg)
        nu, nv = int(proj_2d_shape[0]), int(proj_2d_shape[1])
        if _hacky_fan_beam_mode or _hacky_fdk_mode:
            return (nu, nv)
        width = nv * np.sin(max_abs_rotation_angle) + nu * np.cos(max_abs_rotation_angle)
        e = [int(np.ceil(width)), int(np.ceil(nv))]
        e[0] += (e[0] - nu) % 2
        e[1] += (e[1] - nv) % 2
        return tuple(e)

    def _angles(self, n: int, force_rotation_angle: float | None) -> np.ndarray:
        if force_rotation_angle is None:
            return self.uniform_rng(low=-self._max_abs_rotation_angle, high=self._max_abs_rotation_angle, len=n)
        return force_rotation_angle * np.ones((n,), dtype=np.float32)

    def _filter_chunk(self, v, angles: np.ndarray) -> list:
        """[one filtered tensor (n, nu, nv) float32 per kernel] for a device chunk ``v``."""
        torch = _torch()
        lib = _abi.lib()
        n = v.shape[0]
        eu, ev = self._expanded_2d_shape
        dev = v.device
        if self.hacky_fdk_mode and not np.any(angles):
            # no rotation and no expansion: the cosine weight fused into the row filter
            return [ramp_filter_rows(v, self.physpixwid / self.L, kernel=code, fdk_cosine_weight=True)
                    for code in self.rfilt._codes]
        ang = torch.from_numpy(np.ascontiguousarray(angles, dtype=np.float32)).to(dev)
        buf = torch.empty((n, eu, ev), dtype=torch.float32, device=dev)
        st = _abi.stream_handle(dev)
        _abi.check(lib.ap_fbp_rotate_weight(_abi.array3(v), _abi.ptr(v), _abi.array3(buf), _abi.ptr(buf),
                                            _abi.ptr(ang), float(self.physpixwid / self.L),
                                            float(self._max_abs_rotation_angle), self._mode, st),
                   "fbp_rotate_weight")
        outs = []
        for code in self.rfilt._codes:
            filt = ramp_filter_rows(buf, self.physpixwid / self.L, kernel=code)
            o = torch.empty((n, *self.proj_2d_shape), dtype=torch.float32, device=dev)
            _abi.check(lib.ap_fbp_unrotate(_abi.array3(filt), _abi.ptr(filt), _abi.array3(o), _abi.ptr(o),
                                           _abi.ptr(ang), st), "fbp_unrotate")
            outs.append(o)
        return outs

    def _device_chunk(self, arr, sl: slice):
        torch = _torch()
        if isinstance(arr, torch.Tensor):
            v = arr[sl]
            _abi.require_device(v, "projection array")
            return v if v.dtype == torch.float32 else v.float()
        return torch.from_numpy(np.ascontiguousarray(arr[sl], dtype=np.float32)).cuda()

    def _check_input(self, arr) -> None:
        if tuple(arr.shape[1:]) != self.proj_2d_shape:
            msg = "Wrong input size."
            raise ValueError(msg)

    def filter_bulk_axis1_stream(self, arr, block: bool = True, clear_gpu_cache_regularly: bool = False,
                                 force_rotation_angle: float | None = None):
        """Generator of ([filtered chunk per kernel], event) over chunks of ``projs_per_batch`` projections."""
        torch = _torch()
        del clear_gpu_cache_regularly
        self._check_input(arr)
        n = self.chunk_shape[0]
        for i0 in range(0, arr.shape[0], n):
            sl = slice(i0, min(i0 + n, arr.shape[0]))
            outs = self._filter_chunk(self._device_chunk(arr, sl), self._angles(sl.stop - sl.start,
                                                                                force_rotation_angle))
            ev = torch.cuda.Event()
            ev.record()
            self.event_completed_filter = ev
            if block:
                ev.synchronize()
            yield outs, ev

    def filter_bulk_axis1_concat(self, arr, clear_gpu_cache_regularly: bool = True,
                                 force_rotation_angle: float | None = None, block: bool = True):
        """([filtered array per kernel], event) for the whole input.

        The rotation angles are drawn in projection order, as chunk by chunk,
        so the result does not depend on ``projs_per_batch``; the work runs in
        chunks of up to about 1 GiB of expanded detector data."""
        torch = _torch()
        del clear_gpu_cache_regularly
        self._check_input(arr)
        eu, ev_ = self._expanded_2d_shape
        step = min(65535, max(self.chunk_shape[0], (1 << 28) // max(1, eu * ev_)))   # grid z limit
        dev = arr.device if isinstance(arr, torch.Tensor) else torch.device("cuda")
        odt = arr.dtype if isinstance(arr, torch.Tensor) else torch.float32
        outputs = [torch.empty((arr.shape[0], *self.proj_2d_shape), dtype=odt, device=dev)
                   for _ in range(self.rfilt.num_kerns)]
        for i0 in range(0, arr.shape[0], step):
            sl = slice(i0, min(i0 + step, arr.shape[0]))
            outs = self._filter_chunk(self._device_chunk(arr, sl), self._angles(sl.stop - sl.start,
                                                                                force_rotation_angle))
            for o, out in zip(outs, outputs):
                out[sl].copy_(o)
        ev = torch.cuda.Event()
        ev.record()
        self.event_completed_filter = ev
        if block:
            ev.synchronize()
        return outputs, ev


def fdk_reconstruct(projs, volume, *, R: float, L: float, physpixwid: float,  # noqa: N803
                    bilinear: bool = True, cyl_mask_rad: float | None = None):
    """FDK as in ``scripts/recon_cbfp_cyl_fbp.py`` (fdk geometry): filter,
    fast direct-raycast backprojection with weight (Z/R)^-2, then x 1/(4 pi R V).

    ``projs`` is a ProjectionSet with data; ``volume`` receives the result
    (added into it).
    """
    from . import flatpanel as cbfp  # noqa: PLC0415

    filt = FBPFilterer(L=L, physpixwid=physpixwid, proj_2d_shape=projs.data.shape[-2:],
                       _hacky_fdk_mode=True, limit_zeta=0.0)
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
