"""Pins of the oracle's smoothed total variation (no reference counterpart:
aprict has no TV term; the definition is the differentiable workload's)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc


def test_tv_value_closed_forms():
    # constant volume: every difference is zero
    assert orc.tv_value(np.full((3, 4, 5), 7.0), 0.5) == pytest.approx(60 * 0.5)
    # ramp along x with slope a: (nx - 1) * ny * nz interior terms sqrt(a^2 + eps^2),
    # the last x-layer has no forward difference
    x = np.arange(4.0)[:, None, None] * 2.0 * np.ones((4, 3, 2))
    want = 3 * 3 * 2 * np.sqrt(4.0 + 0.01) + 1 * 3 * 2 * 0.1
    assert orc.tv_value(x, 0.1) == pytest.approx(want)
    # a single bright voxel at the centre of a 3^3 volume: its own term sees
    # three differences of -1, three neighbours one each
    y = np.zeros((3, 3, 3))
    y[1, 1, 1] = 1.0
    want = np.sqrt(3 + 1e-4) + 3 * np.sqrt(1 + 1e-4) + 23 * 1e-2
    assert orc.tv_value(y, 1e-2) == pytest.approx(want)


def test_tv_gradient_is_the_derivative_of_tv_value():
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, (4, 3, 5))
    g = orc.tv_gradient(x, 0.2)
    h = 1e-6
    fd = np.zeros_like(x)
    for idx in np.ndindex(x.shape):
        e = np.zeros_like(x)
        e[idx] = h
        fd[idx] = (orc.tv_value(x + e, 0.2) - orc.tv_value(x - e, 0.2)) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-8)


def test_tv_gradient_sums_to_zero_and_vanishes_on_constants():
    # TV is invariant to adding a constant: the gradient sums to zero
    rng = np.random.default_rng(6)
    x = rng.uniform(0, 1, (5, 4, 3))
    assert abs(orc.tv_gradient(x, 1e-3).sum()) < 1e-9
    assert np.abs(orc.tv_gradient(np.full((3, 3, 3), 2.0), 1e-3)).max() == 0.0
