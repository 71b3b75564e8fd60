"""numpy / scipy restatement of FBPFilterer's filtering chain.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the FBP filter;
the product package never imports it.

Follows src/aprict/cb/ramp_filter.py: the largest rotation angle (:508-530),
the expanded detector (:532-575), the detector weights (:407-449), the
letterbox (:470-500, 689-754) and the per-projection chain of
filter_bulk_axis1_stream (:577-886).  The rotations are the reference's own
calls made with scipy.ndimage.rotate (the CPU twin of the
cupyx.scipy.ndimage.rotate the reference calls, same arguments); the row
filter is the long-double direct convolution of oracle.c.  Everything else
is FP64.
"""

from __future__ import annotations

import numpy as np
import scipy.ndimage

from . import oracle as orc

GBC, FAN, FDK = 0, 1, 2


def max_rotation_angle(nu: int, nv: int, physpixwid: float, L: float) -> float:  # noqa: N803
    """Half the vertical cone angle at the horizontal edge of the detector."""
    return float(np.arctan2(nv * physpixwid / 2, np.sqrt(L**2 + (nu * physpixwid / 2) ** 2)))


def expanded_shape(nu: int, nv: int, zeta: float, mode: int) -> tuple[int, int]:
    """Width nv sin(zeta) + nu cos(zeta) rounded up, keeping the parity of nu."""
    if mode != GBC:
        return nu, nv
    eu = int(np.ceil(nv * np.sin(zeta) + nu * np.cos(zeta)))
    eu += (eu - nu) % 2
    return eu, nv


def weights(eu: int, ev: int, pw: float, mode: int) -> np.ndarray:
    """Detector weight at unit source distance, u along axis 0, v along axis 1."""
    u = pw * (np.arange(eu) - (eu - 1) / 2)[:, None]
    v = pw * (np.arange(ev) - (ev - 1) / 2)[None, :]
    r = np.sqrt(1 + u * u + v * v)
    if mode == FDK:
        return 1 / r
    return 1 / (np.sqrt(1 + v * v) ** 3 * r)


def letterbox(zeta: float, zeta_max: float, L: float, physpixwid: float, ev: int) -> np.ndarray:  # noqa: N803
    """Per-column (v) factor: 1 inside the rotated detector's v range, the
    covered fraction at its two end columns, 0 outside."""
    mid = (ev - 1) / 2
    disc = L / physpixwid * np.sqrt(max((np.cos(np.float32(zeta)) / np.cos(zeta_max)) ** 2 - 1, 0))
    start, stop = mid - disc, mid + disc
    k1, k2 = round(start), round(stop)
    f = np.zeros(ev)
    if k1 == k2:
        f[k1] = stop - start
    else:
        f[k1 + 1:k2] = 1
        f[k1] = (k1 + 0.5) - start
        if k2 < ev:
            f[k2] = stop - (k2 - 0.5)
    return f


def _rotate(img: np.ndarray, zeta: float) -> np.ndarray:
    # the reference's call: angle in degrees, axes (1, 0), order 1, zeros outside
    return scipy.ndimage.rotate(img, angle=float(np.degrees(zeta)), axes=(1, 0), reshape=False, order=1,
                                mode="grid-constant", cval=0.0, prefilter=False)


def fbp_filter(arr: np.ndarray, *, L: float, physpixwid: float, angles, zeta_max: float, mode: int,  # noqa: N803
               kernels=(0,)) -> list[np.ndarray]:
    """The filtered projections, one [V][nu][nv] array per ramp kernel code
    (0 Ram-Lak, 2 Hilbert: the AP_RAMP_* codes of oracle.c)."""
    arr = np.asarray(arr, dtype=np.float64)
    nv_, nu, nv = arr.shape
    eu, ev = expanded_shape(nu, nv, zeta_max, mode)
    pu, pv = (eu - nu) // 2, (ev - nv) // 2
    w = weights(eu, ev, physpixwid / L, mode)
    outs = [np.empty_like(arr) for _ in kernels]
    for p in range(nv_):
        z = float(np.float32(angles[p]))
        padded = np.pad(arr[p], ((pu, pu), (pv, pv)))
        rot = padded if z == 0 else _rotate(padded, -z)
        rot = rot * w
        if mode == GBC:
            rot = rot * letterbox(z, zeta_max, L, physpixwid, ev)[None, :]
        for q, k in enumerate(kernels):
            filt = orc.ramp_filter(rot[None], physpixwid / L, kernel=k)[0]
            back = filt if z == 0 else _rotate(filt, z)
            outs[q][p] = back[pu:pu + nu, pv:pv + nv]
    return outs
