"""Pins the FP64 oracle of the backprojectors (CPU only).

- matched adjoint: transpose of the dense slab-length system matrix and the
  dot-product identity <Ax,y> = <x,A^T y>;
- exact voxel-driven backprojector with aprict's semantics
  (CB_backproject.cu:408-477): equals the matched adjoint at ss=1;
- fast backprojector (:478-691): closed-form weight at the isocentre,
  exact bilinear reproduction of a linear detector image, mask, softening
  and multiplier semantics, and FDK of a ball recovering its density
  (Feldkamp-Davis-Kress 1984).
"""

from __future__ import annotations

import numpy as np
import pytest

from aprict_amd.cb import flatpanel as cbfp
from aprict_amd.volume import forbild
from helpers import chunk_for, pixel_point, random_geometry, source_point
from oracle import oracle as orc
from test_oracle_siddon import dense_system


def test_matched_adjoint_is_transpose_of_dense_matrix():
    rng = np.random.default_rng(21)
    shape, w = (4, 5, 3), 0.6
    pg = random_geometry(rng, 3, 4, 3, R=7.0, L=14.0, pitch=0.8)
    specs = pg.to_proj_spec_array()
    y = rng.uniform(-1, 1, (3, 4, 3))
    cfg = cbfp.ConfigProject(ssu=2, ssv=1, multiply_before_adding=0.7).contents
    ch = chunk_for(shape, w)
    got = orc.backproject_matched(cfg, specs, ch, np.zeros(shape), y, alpha=1.5)
    A = float(np.float32(0.7)) * dense_system(specs, shape, ch, ssu=2, ssv=1)
    np.testing.assert_allclose(got.ravel(), 1.5 * A.T @ y.ravel(), rtol=1e-11, atol=1e-11)
    # and the dot-product identity against the forward oracle
    x = rng.uniform(0, 1, shape)
    ax = orc.project_voxel(cfg, specs, ch, x, np.zeros((3, 4, 3)))
    aty = orc.backproject_matched(cfg, specs, ch, np.zeros(shape), y)
    assert np.vdot(ax, y) == pytest.approx(np.vdot(x, aty), rel=1e-12)


def test_exact_backprojector_equals_matched_adjoint_at_ss1():
    rng = np.random.default_rng(4)
    shape, w = (6, 5, 7), 0.5
    pg = random_geometry(rng, 4, 7, 6, R=9.0, L=18.0, pitch=0.5)
    specs = pg.to_proj_spec_array()
    y = rng.uniform(0, 1, (4, 7, 6))
    ch = chunk_for(shape, w, pos_c=(0.1, -0.2, 0.05))
    exact = orc.backproject(cbfp.ConfigBackproject(fast_approx=False).contents, specs, ch, np.zeros(shape), y)
    matched = orc.backproject_matched(cbfp.ConfigProject().contents, specs, ch, np.zeros(shape), y)
    assert matched.max() > 0.1
    np.testing.assert_allclose(exact, matched, rtol=1e-10, atol=1e-12)


def _circle(views=1, n=16, pitch=1.0, R=100.0, L=200.0):
    return cbfp.ProjGeom.new_circular_scan(num_source_points=views, R=R, L_minus_R=L - R,
                                           pix_widhei_at_isocenter=(pitch * R / L, pitch * R / L),
                                           num_pix_widhei=(n, n))


def test_fast_weight_at_isocentre_is_magnification_squared():
    # voxel centred at the isocentre, constant image c: the value is
    # c * w^3 (L/R)^2 / (|u||v|)  (the detector sees the voxel magnified L/R)
    pg = _circle(views=1, n=16, pitch=0.8)
    specs = pg.to_proj_spec_array()
    w, c = 0.3, 2.0
    cfg = cbfp.ConfigBackproject(fast_approx=True, approx_mode_bilinear_interpolate=True).contents
    ch = chunk_for((1, 1, 1), w)
    got = orc.backproject(cfg, specs, ch, np.zeros((1, 1, 1)), np.full((1, 16, 16), c))
    w32, p32 = float(np.float32(w)), float(np.float32(0.8))   # as stored in the structs
    assert got[0, 0, 0] == pytest.approx(c * w32**3 * 2.0**2 / p32**2, rel=1e-12)


def test_fast_bilinear_reproduces_a_linear_image():
    # put a one-voxel volume on the ray through fractional detector index
    # (fj, fk); a detector image linear in (j, k) is interpolated exactly
    pg = _circle(views=1, n=12, pitch=1.0)
    specs = pg.to_proj_spec_array()
    fj, fk = 4.25, 6.6
    S, Q = source_point(specs[0]), pixel_point(specs[0], fj, fk)
    pos = S + 0.45 * (Q - S)
    jj, kk = np.meshgrid(np.arange(12), np.arange(12), indexing="ij")
    img = (0.5 + 0.3 * jj - 0.2 * kk)[None].astype(np.float64)
    cfg = cbfp.ConfigBackproject(fast_approx=True, approx_mode_direct_raycast=True,
                                 approx_mode_bilinear_interpolate=True).contents
    got = orc.backproject(cfg, specs, chunk_for((1, 1, 1), 0.2, pos_c=pos), np.zeros((1, 1, 1)), img)
    assert got[0, 0, 0] == pytest.approx(0.5 + 0.3 * fj - 0.2 * fk, rel=5e-6)
    cfg_n = cbfp.ConfigBackproject(fast_approx=True, approx_mode_direct_raycast=True).contents
    got_n = orc.backproject(cfg_n, specs, chunk_for((1, 1, 1), 0.2, pos_c=pos), np.zeros((1, 1, 1)), img)
    assert got_n[0, 0, 0] == pytest.approx(0.5 + 0.3 * 4 - 0.2 * 7, rel=1e-6)


def test_mask_multiplier_and_softening_semantics():
    rng = np.random.default_rng(8)
    pg = _circle(views=5, n=20, pitch=1.0)
    specs = pg.to_proj_spec_array()
    shape, w = (9, 9, 4), 1.0
    y = rng.uniform(0, 1, (5, 20, 20))
    ch = chunk_for(shape, w)
    base = orc.backproject(cbfp.ConfigBackproject(fast_approx=True).contents, specs, ch, np.zeros(shape), y)
    masked = orc.backproject(cbfp.ConfigBackproject(fast_approx=True, cyl_mask_rad=2.6).contents, specs, ch,
                             np.zeros(shape), y)
    xs = (np.arange(9) - 4) * w
    r2 = xs[:, None] ** 2 + xs[None, :] ** 2
    inside = np.broadcast_to((r2 <= 2.6**2)[:, :, None], shape)
    np.testing.assert_allclose(masked[inside], base[inside], rtol=1e-13)
    assert np.all(masked[~inside] == 0)
    mult = rng.uniform(0.5, 2.0, shape)
    got = orc.backproject(cbfp.ConfigBackproject(fast_approx=True, apply_summand_multiplier=True).contents,
                          specs, ch, np.ones(shape), y, mult=mult)
    np.testing.assert_allclose(got, 1.0 + mult * base, rtol=1e-13)
    soft = orc.backproject(cbfp.ConfigBackproject(fast_approx=True, soften_det_edges_proportion=0.5).contents,
                           specs, ch, np.zeros(shape), y)
    assert np.all(soft <= base + 1e-12) and soft[4, 4, 2] == pytest.approx(base[4, 4, 2], rel=1e-12)


def test_fdk_of_a_ball_recovers_its_density():
    R, L, views, n, pitch = 100.0, 200.0, 180, 64, 1.0
    pg = _circle(views=views, n=n, pitch=pitch, R=R, L=L)
    specs = pg.to_proj_spec_array()
    ball = forbild.stack([forbild.sphere((0.0, 0.0, 0.0), 5.0, 1.0)])
    cfg_fb = np.array((1.0, False, 1, 1), dtype=__import__("aprict_amd")._abi.config_cb_project_FORBILD)
    sino = orc.project_forbild(cfg_fb, specs, ball, np.zeros((views, n, n)))
    filt = orc.ramp_filter(sino, pitch / L, fdk=True)
    shape, w = (11, 11, 11), 1.0
    cfg = cbfp.ConfigBackproject(fast_approx=True, Z_on_val_factor_power=-2, Z_on_val_factor_val=R,
                                 approx_mode_direct_raycast=True, approx_mode_bilinear_interpolate=True).contents
    rec = orc.backproject(cfg, specs, chunk_for(shape, w), np.zeros(shape), filt)
    rec /= 4 * np.pi * R * views
    centre = rec[4:7, 4:7, 4:7]
    assert abs(centre.mean() - 1.0) < 0.03
    assert abs(rec[0, 0, 5]) < 0.05   # 7 mm from the centre: outside the 5 mm ball
