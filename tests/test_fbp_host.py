"""Host-side logic of FBPFilterer (CPU): the rotation range, the expanded
detector, the letterbox slice and the angle draws, checked against the
oracle's restatement of cb/ramp_filter.py:451-575; and the input checks."""

from __future__ import annotations

import math

import numpy as np
import pytest

from aprict_amd.cb.ramp_filter import FBPFilterer
from oracle import fbp


@pytest.mark.parametrize("shape, p, L", [((40, 34), 0.9, 60.0), ((768, 768), 1.0, 1000.0), ((33, 64), 0.5, 20.0)])
def test_rotation_range_and_expanded_detector(shape, p, L):  # noqa: N803
    f = FBPFilterer(L=L, physpixwid=p, proj_2d_shape=shape)
    assert f.zeta == pytest.approx(fbp.max_rotation_angle(*shape, p, L), rel=1e-15)
    assert f._expanded_2d_shape == fbp.expanded_shape(*shape, f.zeta, fbp.GBC)
    assert (f._expanded_2d_shape[0] - shape[0]) % 2 == 0
    # the rotated detector fits the expanded width
    assert f._expanded_2d_shape[0] >= shape[1] * math.sin(f.zeta) + shape[0] * math.cos(f.zeta)
    for kw in (dict(_hacky_fan_beam_mode=True), dict(_hacky_fdk_mode=True)):
        assert FBPFilterer(L=L, physpixwid=p, proj_2d_shape=shape, **kw)._expanded_2d_shape == shape
    assert FBPFilterer(L=L, physpixwid=p, proj_2d_shape=shape, limit_zeta=0.01).zeta == min(f.zeta, 0.01)


def test_letterbox_slice_matches_oracle():
    f = FBPFilterer(L=60.0, physpixwid=0.9, proj_2d_shape=(40, 34))
    for z in (0.0, 0.03, -0.1, 0.999 * f.zeta):
        sl = f.get_letterbox_slice(np.float32(z))
        lb = fbp.letterbox(z, f.zeta, 60.0, 0.9, f._expanded_2d_shape[1])
        assert sl.stop - sl.start == pytest.approx(lb.sum(), rel=1e-12, abs=1e-12)
    with pytest.raises(ValueError, match="Invalid rotation angle"):
        f.get_letterbox_slice(2 * f.zeta)


def test_angle_draws_and_rng_advance():
    f = FBPFilterer(L=60.0, physpixwid=0.9, proj_2d_shape=(40, 34), rng=np.random.default_rng(5))
    a = f.uniform_rng(-f.zeta, f.zeta, 4)
    rng = np.random.default_rng(5)
    want = np.array([-f.zeta + 2 * f.zeta * rng.random() for _ in range(4)], dtype=np.float32)
    np.testing.assert_array_equal(a, want)
    f.advance_rng(3)
    [rng.random() for _ in range(3)]
    assert f.rng.random() == rng.random()


def test_mode_and_shape_checks():
    with pytest.raises(ValueError, match="both"):
        FBPFilterer(L=60.0, physpixwid=0.9, proj_2d_shape=(40, 34), _hacky_fan_beam_mode=True, _hacky_fdk_mode=True)
    f = FBPFilterer(L=60.0, physpixwid=0.9, proj_2d_shape=(40, 34))
    with pytest.raises(ValueError, match="Wrong input size"):
        f.filter_bulk_axis1_concat(np.zeros((2, 40, 33), np.float32))
