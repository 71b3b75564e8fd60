"""Integrated GBC weights over a perfect cylindrical source locus
(``src/aprict/gbc/weightinteg.py:24-98``).

By symmetry the integral depends only on the point's distance r from the
cylinder's axis (in units of the source radius R); ``ap_gbc_weight_integ``
evaluates it by quadrature over source angle and height.
"""

from __future__ import annotations

import numpy as np

from .. import _abi
from ..cb import flatpanel as cbfp


def backproj_gbc_weight_integ_GPU(r_on_R: np.ndarray, config: cbfp.ConfigBackproject,  # noqa: N802, N803
                                  num_z_samps: int, num_ang_samps: int) -> np.ndarray:
    """Integral of the preset-7 weight times its softening over the source locus,
    for each r / R in the increasing, non-negative 1-D array ``r_on_R`` (float32)."""
    import torch  # noqa: PLC0415

    r_on_R = np.asarray(r_on_R)  # noqa: N806
    if r_on_R.ndim != 1:
        msg = "r_on_R should be a 1-dimensional array."
        raise ValueError(msg)
    dev = torch.device("cuda")
    r = torch.from_numpy(r_on_R.astype(np.float32)).to(dev)
    out = torch.zeros(r.shape, dtype=torch.float64, device=dev)
    cfg = config.contents
    _abi.check(_abi.lib().ap_gbc_weight_integ(_abi.ptr(r), len(r_on_R), _abi.ptr(out), int(num_z_samps),
                                              int(num_ang_samps), _abi.struct_ptr(cfg), _abi.stream_handle(dev)),
               "gbc_weight_integ")
    return out.cpu().numpy().astype(np.float32)


def backproj_gbc_weight_integ_xygrid_GPU(xs: np.ndarray, ys: np.ndarray, R: float,  # noqa: N802, N803
                                         config: cbfp.ConfigBackproject, num_z_samps: int = 6000,
                                         num_ang_samps: int = 6000) -> np.ndarray:
    """The integrated weight on the grid xs x ys: evaluated at 2 max(len(xs), len(ys))
    radii spanning the grid's range of r / R and interpolated linearly."""
    rel = np.sqrt(np.square(xs[:, np.newaxis]) + np.square(ys[np.newaxis, :])) / R
    numsamps = 2 * max(len(xs), len(ys))
    r_in = np.linspace(rel.min(), rel.max(), endpoint=True, num=numsamps)
    integ = backproj_gbc_weight_integ_GPU(r_in, config, num_z_samps, num_ang_samps)
    return np.interp(rel, r_in, integ)
