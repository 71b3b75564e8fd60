"""Preconditioned SIRT on GPU-resident data (drop-in for ``aprict.iterators.sirt_vram``).

Same algorithm as ``src/aprict/iterators/sirt_vram.py:62-149``:

- preconditioner ``M = s * diag(B A 1_cyl)^-1`` built from a hard-masked
  constant volume, the mean-normalised forward projection and the
  (by default fast) backprojection, inverted with the soft-zero guard
  ``1/(v + 0.1 exp(-v/0.1))``, then soft-masked to the cylinder;
- ``iterate()``: ``x += M (.) B (b - A x)`` with the cylinder-masked,
  edge-softened backprojector ``B``.

Each iteration is one fused library call (``ap_sirt_iterate``): the residual
is written once by the projector epilogue (no copy of ``b``, no
accumulate pass) and the preconditioner multiply happens in the
backprojector epilogue.
"""

from __future__ import annotations

import dataclasses

import numpy as np

from .. import _abi
from ..cb import flatpanel as cbfp
from ..volume import Volume


def _torch():
    import torch  # noqa: PLC0415

    return torch


def _mean(t) -> float:
    """Mean of a 3-D tensor: an FP64 dot product with a broadcast 1 in the library."""
    torch = _torch()
    one = torch.ones(1, dtype=t.dtype, device=t.device).expand(t.shape)
    out = torch.empty(1, dtype=torch.float64, device=t.device)
    _abi.check(_abi.lib().ap_array_dot(_abi.array3(t), _abi.ptr(t), _abi.array3(one), _abi.ptr(one),
                                       _abi.ptr(out), _abi.stream_handle(t.device)), "mean")
    return float(out.item()) / max(t.numel(), 1)


class SIRT_flatpanel:  # noqa: N801
    """An iterable SIRT reconstructor."""

    @dataclasses.dataclass(frozen=True, kw_only=True)
    class Config:
        fast_backproject: bool = True
        soft_from_cylinder_relrad: float | None = 0.95
        bp_soft_from_detedge_relrad: float | None = 0.95

    def __init__(self, config: SIRT_flatpanel.Config, volume: Volume, projs: cbfp.ProjectionSet) -> None:
        torch = _torch()
        self._config = config
        self._vol = volume
        self._meas = projs
        if projs.data is None:
            msg = "The provided ProjectionSet object has no data."
            raise ValueError(msg)
        sep = 0 if config.bp_soft_from_detedge_relrad is None else 1 - config.bp_soft_from_detedge_relrad
        cyl = min(*volume.shape[:2]) / 2 * volume.voxwid
        self._config_bp = cbfp.ConfigBackproject(fast_approx=config.fast_backproject, cyl_mask_rad=cyl,
                                                 apply_summand_multiplier=True,
                                                 soften_det_edges_proportion=sep)
        self._config_fp = cbfp.ConfigProject()
        # preconditioner M = diag(A^T A 1)^-1 with masking (sirt_vram.py:82-116)
        M = volume.zeroed_copy()  # noqa: N806
        p = projs.zeroed_copy()
        scale = 1.0 / float(np.mean(volume.shape))
        M.fill(scale)
        M.soften_to_xy_cylinder(1.0)
        cbfp.project(self._config_fp, M, p)
        pf = p._flat_data()
        meanval = _mean(pf)
        _abi.check(_abi.lib().ap_array_scale(_abi.array3(pf), _abi.ptr(pf), scale / meanval,
                                             _abi.stream_handle(pf.device)), "normalise")
        scale /= meanval / scale
        M.fill(0.0)
        bp_nomult = cbfp.ConfigBackproject(fast_approx=config.fast_backproject, cyl_mask_rad=cyl,
                                           apply_summand_multiplier=False, soften_det_edges_proportion=sep)
        cbfp.backproject(bp_nomult, M, p)
        M.invert(ignore_zeros=True, softmax_epsilon=0.1)
        M.multiply(scale)
        if config.soft_from_cylinder_relrad is not None:
            M.soften_to_xy_cylinder(config.soft_from_cylinder_relrad)
        self._preconditioner_volume = M
        self._proj_scratch_space = p
        # upload the view table now: iterate() then makes no host->device copy
        # and can be recorded into a HIP graph
        projs.specs_device()
        torch.cuda.synchronize(volume.device)

    @property
    def preconditioner(self) -> Volume:
        return self._preconditioner_volume

    def iterate(self) -> None:
        """x += M (.) B (b - A x), one fused library call."""
        x = self._vol.data
        M = self._preconditioner_volume.data  # noqa: N806
        b = self._meas._flat_data()
        r = self._proj_scratch_space._flat_data()
        cfg = self._config_bp.contents
        chunk = self._vol.chunk_spec()
        _abi.check(_abi.lib().ap_sirt_iterate(
            _abi.struct_ptr(cfg), _abi.ptr(self._meas.specs_device()), b.shape[0], _abi.struct_ptr(chunk),
            _abi.array3(x), _abi.ptr(x), _abi.array3(M), _abi.ptr(M), _abi.array3(b), _abi.ptr(b),
            _abi.array3(r), _abi.ptr(r), _abi.stream_handle(x.device)), "sirt_iterate")

    @property
    def volume(self) -> Volume:
        return self._vol

    @property
    def residual(self):
        """``b - A x`` as of the start of the last iteration."""
        return self._proj_scratch_space.data
