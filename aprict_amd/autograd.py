"""Differentiable projector: ``torch.autograd.Function`` around the Siddon operator.

``forward`` is ``ap_project_voxel`` (y = A x) and ``backward`` is its exact
transpose ``ap_backproject_matched`` (grad_x = A^T grad_y), so gradients of
any loss through the projector are exact up to FP32 rounding.  No reference
counterpart: aprict's operators are not differentiable; its exact
backprojector (CB_backproject.cu:408-477) is this adjoint at ssu=ssv=1.
"""

from __future__ import annotations

import numpy as np

from . import _abi
from .cb import flatpanel as cbfp


def _torch():
    import torch  # noqa: PLC0415

    return torch


class Projector:
    """Fixed geometry + config + voxel grid; ``A(x)`` / ``AT(y)`` on raw tensors."""

    def __init__(self, pg: cbfp.ProjGeom, *, vol_shape, voxwid: float, pos_c=(0.0, 0.0, 0.0),
                 config: cbfp.ConfigProject | None = None, device=None) -> None:
        torch = _torch()
        self.pg = pg
        self.vol_shape = tuple(int(s) for s in vol_shape)
        self.voxwid = float(voxwid)
        self.pos_c = np.asarray(pos_c, dtype=np.float64)
        self.config = config or cbfp.ConfigProject()
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self._ps = cbfp.ProjectionSet(pg, device=self.device)
        self._specs = self._ps.specs_device()
        self.sino_shape = (pg.num_projs, int(pg.nu.ravel()[0]), int(pg.nv.ravel()[0]))
        texpos = self.pos_c - self.voxwid * np.array(self.vol_shape, np.float64) / 2
        self._chunk = np.array((self.voxwid, (0, 0, 0), self.vol_shape, tuple(texpos)),
                               dtype=_abi.vol_tex_chunk_spec)
        self._cfg = self.config.contents
        # the same operator with the sign of multiply_before_adding flipped:
        # A_residual's fused epilogue writes b - A x
        self._cfg_neg = self.config.contents
        self._cfg_neg["multiply_before_adding"] = -self._cfg_neg["multiply_before_adding"]

    def A(self, x, out=None, b=None):  # noqa: N802
        """out = (b or 0) + A x (out float32, allocated if None)."""
        torch = _torch()
        if tuple(x.shape) != self.vol_shape:
            msg = f"volume shape {tuple(x.shape)} != {self.vol_shape}"
            raise ValueError(msg)
        _abi.require_device(x, "volume")
        if out is None:
            out = torch.zeros(self.sino_shape, dtype=torch.float32, device=x.device)
        _abi.check(_abi.lib().ap_project_voxel(
            _abi.struct_ptr(self._cfg), _abi.ptr(self._specs), self.sino_shape[0], _abi.struct_ptr(self._chunk),
            _abi.array3(x), _abi.ptr(x), _abi.array3(out), _abi.ptr(out),
            _abi.array3(b) if b is not None else None, _abi.ptr(b) if b is not None else None,
            _abi.stream_handle(x.device)), "A")
        return out

    def A_residual(self, x, b, out=None):  # noqa: N802
        """out = b - A x in one pass (the projector's fused residual epilogue)."""
        torch = _torch()
        if tuple(x.shape) != self.vol_shape or tuple(b.shape) != self.sino_shape:
            msg = f"shapes {tuple(x.shape)}, {tuple(b.shape)} != {self.vol_shape}, {self.sino_shape}"
            raise ValueError(msg)
        _abi.require_device(x, "volume")
        _abi.require_device(b, "measurements")
        if out is None:
            out = torch.empty(self.sino_shape, dtype=torch.float32, device=x.device)
        _abi.check(_abi.lib().ap_project_voxel(
            _abi.struct_ptr(self._cfg_neg), _abi.ptr(self._specs), self.sino_shape[0], _abi.struct_ptr(self._chunk),
            _abi.array3(x), _abi.ptr(x), _abi.array3(out), _abi.ptr(out), _abi.array3(b), _abi.ptr(b),
            _abi.stream_handle(x.device)), "A_residual")
        return out

    def AT(self, y, out=None, alpha: float = 1.0):  # noqa: N802
        """out += alpha A^T y (out float32 zeros if None)."""
        torch = _torch()
        if tuple(y.shape) != self.sino_shape:
            msg = f"sinogram shape {tuple(y.shape)} != {self.sino_shape}"
            raise ValueError(msg)
        _abi.require_device(y, "sinogram")
        if out is None:
            out = torch.zeros(self.vol_shape, dtype=torch.float32, device=y.device)
        ws, wsb = _abi.matched_workspace(self._cfg, self.sino_shape[0], y)
        _abi.check(_abi.lib().ap_backproject_matched(
            _abi.struct_ptr(self._cfg), _abi.ptr(self._specs), self.sino_shape[0], _abi.struct_ptr(self._chunk),
            _abi.array3(out), _abi.ptr(out), None, None, _abi.array3(y), _abi.ptr(y), float(alpha), ws, wsb,
            _abi.stream_handle(y.device)), "AT")
        return out


def _make_function():
    torch = _torch()

    class _ProjectFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, projector):
            ctx.projector = projector
            return projector.A(x.detach())

        @staticmethod
        def backward(ctx, grad_y):
            g = ctx.projector.AT(grad_y.detach().float().contiguous())
            return g, None

    return _ProjectFn


_FN = None


def project(x, projector: Projector):
    """Differentiable ``A x`` (float32 sinogram)."""
    global _FN  # noqa: PLW0603
    if _FN is None:
        _FN = _make_function()
    return _FN.apply(x, projector)


def tv_value(x, eps: float = 1e-3) -> float:
    """sum_i sqrt(|grad_fwd x_i|^2 + eps^2) accumulated in FP64 (HIP kernel)."""
    torch = _torch()
    out = torch.empty(1, dtype=torch.float64, device=x.device)
    _abi.check(_abi.lib().ap_tv_value(_abi.array3(x), _abi.ptr(x), float(eps), _abi.ptr(out),
                                      _abi.stream_handle(x.device)), "tv_value")
    return float(out.item())


def tv_gradient(x, weight: float, eps: float = 1e-3, out=None):
    """out += weight * grad TV_eps(x) (isotropic, forward differences), HIP kernel."""
    torch = _torch()
    if out is None:
        out = torch.zeros_like(x, dtype=torch.float32)
    _abi.check(_abi.lib().ap_tv_gradient(_abi.array3(x), _abi.ptr(x), _abi.array3(out), _abi.ptr(out),
                                         float(weight), float(eps), _abi.stream_handle(x.device)), "tv")
    return out
