"""Landweber with ordered subsets (drop-in for ``aprict.iterators.landweber_vram``).

Same algorithm as ``src/aprict/iterators/landweber_vram.py:73-260``: a
per-subset step ``accel/lambda`` with ``lambda`` the Rayleigh quotient of
``B A`` from ``rayleigh_quotient_num_iterations`` power iterations on a random
positive volume; ``iterate()`` does, per subset,
``r = z (.) (b - A x)``, ``x += (accel/lambda) B r``.  The residual comes from
the fused projector epilogue; masks and scaling are single element-wise
library launches; norms are FP64 sums in the library.
"""

from __future__ import annotations

import dataclasses
import logging

import numpy as np

from .. import _abi
from ..cb import flatpanel as cbfp
from ..volume import Volume


def _torch():
    import torch  # noqa: PLC0415

    return torch


def _apply_mask(projs: cbfp.ProjectionSet, zm) -> None:
    """projs.data *= zm (zm broadcast to the data's shape), one library call."""
    torch = _torch()
    data = projs.data
    m = torch.as_tensor(zm, device=data.device).to(data.dtype).expand(data.shape)
    flat = projs._flat_data()
    try:
        m3 = m.view(flat.shape)
    except RuntimeError:
        m3 = m.contiguous().view(flat.shape)
    _abi.check(_abi.lib().ap_array_mul(_abi.array3(flat), _abi.ptr(flat), _abi.array3(m3), _abi.ptr(m3),
                                       _abi.stream_handle(data.device)), "mask")


def _sq_norm(t) -> float:
    """sum t^2 accumulated in FP64 on the GPU."""
    torch = _torch()
    out = torch.empty(1, dtype=torch.float64, device=t.device)
    _abi.check(_abi.lib().ap_array_dot(_abi.array3(t), _abi.ptr(t), _abi.array3(t), _abi.ptr(t), _abi.ptr(out),
                                       _abi.stream_handle(t.device)), "norm")
    return float(out.item())


class Landweber_flatpanel:  # noqa: N801
    """An iterable Landweber reconstructor."""

    @dataclasses.dataclass(frozen=True, kw_only=True)
    class Config:
        rayleigh_quotient_num_iterations: int = 10
        acceleration_factor: float = 1
        fast_backproject: bool = True
        soft_from_cylinder_relrad: float | None = 0.95
        bp_soft_from_detedge_relrad: float | None = 0.95

    def __init__(self, config: Landweber_flatpanel.Config, volume: Volume, projs, *,
                 proj_zero_masks=None, logger: logging.Logger | None = None) -> None:
        self._logger = logger if logger is not None else logging.getLogger()
        self._config = config
        self._vol = volume
        self._meas = projs if isinstance(projs, list) else [projs]
        if proj_zero_masks is None:
            self._proj_zero_masks = [None for _ in self._meas]
        elif isinstance(projs, list):
            if not isinstance(proj_zero_masks, list):
                msg = "The input measurements were a list, but only a single mask was provided."
                raise ValueError(msg)
            self._proj_zero_masks = proj_zero_masks
        else:
            if isinstance(proj_zero_masks, list):
                msg = "The input measurements were not a list, but a list of masks was provided."
                raise ValueError(msg)
            self._proj_zero_masks = [proj_zero_masks]
        sep = 0 if config.bp_soft_from_detedge_relrad is None else 1 - config.bp_soft_from_detedge_relrad
        self._config_fp = cbfp.ConfigProject()
        self._config_bp = cbfp.ConfigBackproject(fast_approx=config.fast_backproject,
                                                 cyl_mask_rad=min(*volume.shape[:2]) / 2 * volume.voxwid,
                                                 apply_summand_multiplier=False,
                                                 soften_det_edges_proportion=sep)
        for p in self._meas:
            if p.data is None:
                msg = "A provided ProjectionSet object has no data."
                raise ValueError(msg)
        self._proj_scratch_space = [p.zeroed_copy() for p in self._meas]
        self._rayleigh_quotients = self.calculate_rayleigh_quotients(
            num_iterations=config.rayleigh_quotient_num_iterations)
        for p, zm in zip(self._meas, self._proj_zero_masks):
            if zm is not None:
                _apply_mask(p, zm)

    @property
    def rayleigh_quotients(self) -> list[float]:
        return list(self._rayleigh_quotients)

    def calculate_rayleigh_quotients(self, num_iterations: int) -> list[float]:
        """Rayleigh quotient of B A per subset by the power method (landweber_vram.py:127-209).

        Per subset: a random positive volume v, then ``num_iterations`` times
        v <- v / |v|, rq = |A v|^2, v <- B A v.  Norms are FP64 sums on the GPU.
        """
        use_existing = self._vol.norm == 0
        scratch = self._vol if use_existing else self._vol.zeroed_copy()
        scale = 1.0 / float(np.mean(scratch.shape))
        out = []
        for sp in self._proj_scratch_space:
            v = np.abs(np.random.randn(*scratch.shape) * scale).astype(np.float32)  # noqa: NPY002
            scratch.copy_from(v)
            norm = scratch.norm
            rq = 0.0
            for _ in range(num_iterations):
                scratch.multiply(1.0 / norm)
                sp.data.zero_()
                cbfp.project(self._config_fp, scratch, sp)
                rq = _sq_norm(sp._flat_data())
                scratch.fill(0.0)
                cbfp.backproject(self._config_bp, scratch, sp)
                norm = scratch.norm
            out.append(rq)
        if use_existing:
            self._vol.fill(0.0)
        return out

    def iterate(self) -> None:
        """One pass over the ordered subsets (landweber_vram.py:221-260)."""
        lib = _abi.lib()
        x = self._vol.data
        chunk = self._vol.chunk_spec()
        fp = np.array((1.0, -1.0, 1, 1), dtype=_abi.config_cb_project)
        for meas, scratch, rq, zm in zip(self._meas, self._proj_scratch_space, self._rayleigh_quotients,
                                         self._proj_zero_masks):
            b = meas._flat_data()
            r = scratch._flat_data()
            # r = b - A x in the projector epilogue (b and r with their own strides)
            _abi.check(lib.ap_project_voxel(_abi.struct_ptr(fp), _abi.ptr(meas.specs_device()), b.shape[0],
                                            _abi.struct_ptr(chunk), _abi.array3(x), _abi.ptr(x),
                                            _abi.array3(r), _abi.ptr(r), _abi.array3(b), _abi.ptr(b),
                                            _abi.stream_handle(x.device)), "residual")
            if zm is not None:
                _apply_mask(scratch, zm)
            _abi.check(lib.ap_array_scale(_abi.array3(r), _abi.ptr(r), self._config.acceleration_factor / rq,
                                          _abi.stream_handle(x.device)), "scale")
            cbfp.backproject(self._config_bp, self._vol, scratch)

    @property
    def volume(self) -> Volume:
        return self._vol
