"""CPU checks of the oracle's GBC / FBP pieces (what pins them):

- the preset-7 weight at points where the geometry gives it in closed form
  (on the source axis, along it, sideways, above the plane, inside the
  elevation softening band where the window is exactly 1/2);
- presets 1-4 and 6 of the direct ray-casting branch on one voxel and one
  view, the horizontal geometry (distance ratio, cosine, offsets) worked out
  by hand from the angle the voxel makes with the central ray;
- the weight integral against closed forms (on the axis, the elevation
  integral of cos^3 is z / sqrt(1 + z^2)) and against an independent 2-D
  adaptive quadrature of the pinned point weight over the whole cylinder;
- the letterbox as a partition of the covered range (weights in [0, 1]
  summing to its length), and the FDK chain at angle 0 against the oracle's
  fused cosine-weighted filter.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import scipy.integrate

from aprict_amd import phantoms as ph
from aprict_amd.cb import flatpanel as cbfp
from helpers import chunk_for
from oracle import fbp
from oracle import oracle as orc

R = 40.0


def gbc_cfg(va=1.2, vb=0.1, ha=1.4, hb=0.1, R_=R):  # noqa: N803
    return cbfp.ConfigBackproject(fast_approx=True, backproj_weighting_preset=7, Z_on_val_factor_val=R_,
                                  soft_vartheta_v=va, soft_vartheta_v_softspan=vb, soft_vartheta_h=ha,
                                  soft_vartheta_h_softspan=hb).contents


def test_gbc_weight_closed_forms():
    cfg = gbc_cfg()
    S = (R, 0.0, 0.0)
    # on the axis, in the source plane: theta = 0, c = 1, d = 0
    assert orc.gbc_weight(cfg, S, (0, 0, 0)) == pytest.approx(1.0, rel=1e-14)
    # along the central ray at x: ell = 1 - x/R = c, d = x/R, w = 1 / (1 + (x/R)^2)
    assert orc.gbc_weight(cfg, S, (R / 2, 0, 0)) == pytest.approx(1 / 1.25, rel=1e-14)
    # beyond the axis: ell = 1 + x/R, c = 1, d = -x/R
    assert orc.gbc_weight(cfg, S, (-R / 2, 0, 0)) == pytest.approx(1 / 1.25, rel=1e-14)
    # sideways by R: horizontal angle pi/4, ell = sqrt 2, c = 1/sqrt 2, d = -1/sqrt 2
    assert orc.gbc_weight(cfg, S, (0, R, 0)) == pytest.approx(1 / math.sqrt(2), rel=1e-14)
    # above the plane: elevation theta with tan(theta) = 0.3, w = cos^3(theta)
    assert orc.gbc_weight(cfg, S, (0, 0, 0.3 * R)) == pytest.approx(1.09**-1.5, rel=1e-14)
    # elevation inside the band [a - b, a] where sin(theta) is the mean of sin(a - b) and sin(a): window 1/2
    a, b = float(np.float32(0.6)), float(np.float32(0.2))   # the config holds FP32 angles
    th = math.asin((math.sin(a - b) + math.sin(a)) / 2)
    cfg2 = gbc_cfg(va=a, vb=b)
    assert orc.gbc_weight(cfg2, S, (0, 0, R * math.tan(th))) == pytest.approx(0.5 * math.cos(th) ** 3, rel=1e-12)
    # outside the elevation window, and outside the horizontal one (pi/4 > 0.7)
    assert orc.gbc_weight(cfg2, S, (0, 0, R * math.tan(0.61))) == 0.0
    assert orc.gbc_weight(gbc_cfg(ha=0.7), S, (0, R, 0)) == 0.0


@pytest.mark.parametrize("preset", [0, 1, 2, 3, 4, 5, 6])
def test_preset_weights_on_one_voxel(preset):
    # one view of a circular scan (source at (R, 0, 0), central ray along -x),
    # a detector of ones read at the nearest pixel: the backprojection of a
    # voxel is its weight.  Voxel at (0, y, 0): horizontal angle phi with
    # tan(phi) = y / R, so ell = 1 / cos(phi), c = cos(phi), d = c - ell,
    # Z / R = 1, the offset across the source direction y / R and along it 0.
    pg = ph.circular_geometry(1, 160, 8, 1.0, R=R, L=2 * R)   # 160 mm wide: holds the voxel's shadow at 40 mm
    y = np.ones((1, 160, 8))
    cfg = cbfp.ConfigBackproject(fast_approx=True, approx_mode_direct_raycast=True, Z_on_val_factor_val=R,
                                 backproj_weighting_preset=preset).contents
    for frac in (0.0, 0.5):
        yy = frac * R
        got = orc.backproject(cfg, pg.to_proj_spec_array(), chunk_for((1, 1, 1), 0.1, (0.0, yy, 0.0)),
                              np.zeros((1, 1, 1)), y)[0, 0, 0]
        phi = math.atan(frac)
        ell, c = 1 / math.cos(phi), math.cos(phi)
        d, vy, q = c - ell, frac, (c - ell) / c
        f = (1 - q * q) / (1 + q * q)
        want = {1: 1 / (c * c * (c * c + d * d)), 2: -2 * f * f * (-vy) / (ell * ell), 3: 1 / (c * c),
                4: vy, 6: -2 * f * f / (1 - q) * (-vy)}.get(preset, 1.0)
        # the geometry and the chunk position are FP32 structs: 1e-6 relative
        assert got == pytest.approx(want, rel=1e-6, abs=1e-12), (preset, frac)


def test_weight_integral_on_axis_closed_form():
    # r = 0: c = 1, d = 0 for every source angle, w = cos^3(theta) s(theta), tan(theta) = z;
    # without softening in range the z-integral of (1 + z^2)^(-3/2) is z / sqrt(1 + z^2)
    va = 1.2
    cfg = gbc_cfg(va=va, vb=1e-6)
    zmax = math.tan(va)
    nz = 8000
    got = orc.gbc_weight_integ([0.0], cfg, nz, 3)[0]
    want = 4 * math.pi * zmax / math.sqrt(1 + zmax * zmax)
    assert got == pytest.approx(want, rel=2e-3)


def test_weight_integral_matches_adaptive_quadrature():
    # off the axis: 2-D adaptive quadrature of the (pinned) point weight over
    # source angle t in [0, 2 pi) and height z in [-zmax, zmax]
    cfg = gbc_cfg(va=0.5, vb=0.15, ha=1.2, hb=0.3, R_=1.0)
    r = 0.35
    zmax = math.tan(0.5) * (1 + r)

    def w(z, t):
        return orc.gbc_weight(cfg, (math.cos(t), math.sin(t), z), (r, 0.0, 0.0))

    want, _ = scipy.integrate.dblquad(w, 0.0, 2 * math.pi, -zmax, zmax, epsabs=1e-6, epsrel=1e-6)
    got = orc.gbc_weight_integ([0.0, r], cfg, 3000, 1500)[1]
    assert got == pytest.approx(want, rel=2e-3)


@pytest.mark.parametrize("zeta", [0.0, 0.05, -0.12, 0.2])
def test_letterbox_partitions_the_covered_range(zeta):
    L, p, ev = 30.0, 1.0, 16  # noqa: N806
    zmax = fbp.max_rotation_angle(20, ev, p, L)
    f = fbp.letterbox(zeta, zmax, L, p, ev)
    # half-length of the covered v range; cos of the FP32 angle rounded to FP32, as the reference takes it
    disc = L / p * math.sqrt(max((float(np.cos(np.float32(zeta))) / math.cos(zmax)) ** 2 - 1, 0))
    assert np.all((f >= 0) & (f <= 1))
    assert f.sum() == pytest.approx(2 * disc, rel=1e-12)


def test_fbp_chain_fdk_at_zero_angle_equals_fused_filter():
    rng = np.random.default_rng(3)
    a = rng.uniform(-1, 1, (2, 18, 12))
    L, p = 25.0, 0.8  # noqa: N806
    got = fbp.fbp_filter(a, L=L, physpixwid=p, angles=[0.0, 0.0], zeta_max=0.1, mode=fbp.FDK)[0]
    want = orc.ramp_filter(a, p / L, kernel=0, fdk=True)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
