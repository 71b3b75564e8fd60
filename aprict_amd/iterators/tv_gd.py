"""Gradient descent on ``||A x - b||^2 + beta * TV_eps(x)`` through the
differentiable Siddon projector.

aprict has no such solver (its iterators are SIRT and Landweber, and its
operators are not differentiable); this is the workload that the projector's
matched adjoint serves: the gradient of the data term is ``2 A^T (A x - b)``
with ``A^T`` the exact transpose of ``A`` (``ap_backproject_matched``, the
backward of ``aprict_amd.autograd.project``), and the regulariser is the
smoothed isotropic total variation ``sum sqrt(|D x|^2 + eps^2)`` (forward
differences ``D``).  One ``iterate()`` is four library calls and no host
synchronisation:

    r  = b - A x                       (projector with the fused residual)
    g  = -tau beta grad TV_eps(x)      (fill + ap_tv_gradient)
    g += 2 tau A^T r                   (ap_backproject_matched, alpha = 2 tau)
    x += g                             (ap_array_axpby; x may be stored bf16/fp16)

with the fixed step ``tau = 1 / L``, ``L = 2 * 1.05 * lambda + 12 beta / eps``:
``lambda`` is the largest eigenvalue of ``A^T A`` from a power iteration
(a Rayleigh quotient, hence the 5 % margin) and ``12 / eps`` bounds the
Hessian of TV_eps (``||D||^2 <= 12`` in 3-D).
"""

from __future__ import annotations

import numpy as np

from .. import _abi
from ..autograd import Projector, tv_gradient, tv_value


def _torch():
    import torch  # noqa: PLC0415

    return torch


def _sq_norm(t) -> float:
    torch = _torch()
    out = torch.empty(1, dtype=torch.float64, device=t.device)
    _abi.check(_abi.lib().ap_array_dot(_abi.array3(t), _abi.ptr(t), _abi.array3(t), _abi.ptr(t), _abi.ptr(out),
                                       _abi.stream_handle(t.device)), "norm")
    return float(out.item())


def power_iteration(projector: Projector, iterations: int = 10, seed: int = 0) -> float:
    """Largest eigenvalue of A^T A: Rayleigh quotients ||A v||^2 of unit v <- A^T A v."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    v = torch.from_numpy(rng.uniform(0.0, 1.0, projector.vol_shape).astype(np.float32)).to(projector.device)
    y = torch.empty(projector.sino_shape, dtype=torch.float32, device=projector.device)
    lam = 0.0
    for _ in range(iterations):
        n = np.sqrt(_sq_norm(v))
        if n == 0:
            return 0.0
        _abi.check(_abi.lib().ap_array_scale(_abi.array3(v), _abi.ptr(v), 1.0 / n,
                                             _abi.stream_handle(v.device)), "scale")
        y.zero_()
        projector.A(v, out=y)
        lam = _sq_norm(y)
        v.zero_()
        projector.AT(y, out=v)
    return lam


class TVGradientDescent:
    """``x <- x - tau (2 A^T (A x - b) + beta grad TV_eps(x))`` from ``x0`` (zeros)."""

    def __init__(self, projector: Projector, b, *, beta: float = 1e-3, eps: float = 1e-3, x0=None,
                 dtype=None, lam: float | None = None, power_iterations: int = 10) -> None:
        torch = _torch()
        self.P = projector
        if tuple(b.shape) != projector.sino_shape:
            msg = f"measurements {tuple(b.shape)} != {projector.sino_shape}"
            raise ValueError(msg)
        _abi.require_device(b, "measurements")
        self.b = b if b.dtype == torch.float32 else b.float()
        dtype = dtype or (x0.dtype if x0 is not None else torch.float32)
        self.x = (x0.to(dtype).clone() if x0 is not None
                  else torch.zeros(projector.vol_shape, dtype=dtype, device=projector.device))
        self.r = torch.empty(projector.sino_shape, dtype=torch.float32, device=projector.device)
        self.g = torch.empty(projector.vol_shape, dtype=torch.float32, device=projector.device)
        self.beta = float(beta)
        self.eps = float(eps)
        self.lam = float(lam) if lam is not None else power_iteration(projector, power_iterations)
        self.lipschitz = 2.0 * 1.05 * self.lam + 12.0 * self.beta / self.eps
        self.tau = 1.0 / self.lipschitz
        self.iterations = 0

    def iterate(self) -> None:
        s = _abi.stream_handle(self.x.device)
        lib = _abi.lib()
        self.P.A_residual(self.x, self.b, out=self.r)
        gd, xd = _abi.array3(self.g), _abi.array3(self.x)
        _abi.check(lib.ap_array_fill(gd, _abi.ptr(self.g), 0.0, s), "fill")
        if self.beta != 0.0:
            tv_gradient(self.x, -self.tau * self.beta, self.eps, out=self.g)
        self.P.AT(self.r, out=self.g, alpha=2.0 * self.tau)
        _abi.check(lib.ap_array_axpby(xd, _abi.ptr(self.x), 1.0, gd, _abi.ptr(self.g), 1.0, s), "update")
        self.iterations += 1

    def objective(self) -> tuple[float, float]:
        """(||A x - b||^2, beta TV_eps(x)) at the current x, FP64 sums."""
        self.P.A_residual(self.x, self.b, out=self.r)
        return _sq_norm(self.r), self.beta * tv_value(self.x, self.eps)
