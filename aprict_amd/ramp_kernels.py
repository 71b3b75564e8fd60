"""Ramp-filter taps (host side), same definitions as ``src/aprict/ramp_kernels.py:17-94``.

Each function returns the ``n`` (odd) taps centred on offset 0 in float64;
the GPU filter evaluates the same taps in closed form.
"""

from __future__ import annotations

import numpy as np


def _offsets(n: int) -> np.ndarray:
    if n % 2 == 0:
        msg = "Window length must be odd."
        raise ValueError(msg)
    h = (n - 1) // 2
    return np.arange(-h, h + 1, dtype=np.float64)


def ramp_kernel(n: int, physpixsize: float | None = None) -> np.ndarray:
    """Band-limited ramp with the sinc(f) pixel correction: 8 / (1 - 4 k^2) / p."""
    k = _offsets(n)
    return 8.0 / (1.0 - 4.0 * k * k) / (physpixsize or 1)


def ramp_kernel_ramlak(n: int, physpixsize: float | None = None) -> np.ndarray:
    """Ram-Lak: pi^2/p at k=0, -4/(k^2 p) at odd k, 0 at even k != 0."""
    k = _offsets(n)
    out = np.zeros_like(k)
    out[k == 0] = np.pi**2
    odd = (np.abs(k) % 2) == 1
    out[odd] = -4.0 / (k[odd] ** 2)
    return out / (physpixsize or 1)


def hilbert_transform_kernel(n: int, physpixsize: float | None = None) -> np.ndarray:
    """2/k off the centre, 0 at k=0 (pixel size has no effect)."""
    del physpixsize
    k = _offsets(n)
    out = np.zeros_like(k)
    nz = k != 0
    out[nz] = 2.0 / k[nz]
    return out


def hilbert_transform_kernel_halfoffset(n: int, physpixsize: float | None = None) -> np.ndarray:
    """2/(k + 1/2)."""
    del physpixsize
    k = _offsets(n)
    return 2.0 / (k + 0.5)
