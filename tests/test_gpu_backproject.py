"""GPU parity of the adjoint and the backprojectors with the FP64 oracle.

- matched adjoint (ap_backproject_matched): vs the oracle transpose
  (relative L2 <= 1e-5) and the dot-product test <Ax,y> = <x,A^T y>
  (relative difference < 1e-6, products accumulated in FP64);
- reference backprojector (ap_backproject), exact and fast branches with
  sub-voxel sampling, Z-powers, mask, edge softening and multiplier
  (relative L2 <= 1e-5; the fast branch's FP32 projective map allows
  <= 2e-5 where bilinear weights amplify coordinate rounding).
"""

from __future__ import annotations

import numpy as np
import pytest

from aprict_amd import phantoms as ph
from aprict_amd.cb import flatpanel as cbfp
from aprict_amd.volume import Volume
from helpers import chunk_for, random_geometry, rel_l2
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def to_gpu(a, dtype=None):
    import torch

    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()
    return t if dtype is None else t.to(dtype)


def test_matched_adjoint_matches_oracle(device):
    rng = np.random.default_rng(5)
    shape, w, pc = (36, 30, 28), 0.6, (0.2, 0.1, -0.3)
    pg = random_geometry(rng, 6, 41, 37, R=40.0, L=80.0, pitch=0.9)
    y = rng.uniform(-1, 1, (6, 41, 37)).astype(np.float32)
    mult = rng.uniform(0.5, 1.5, shape).astype(np.float32)
    cfg = cbfp.ConfigProject(ssu=2, ssv=2, multiply_before_adding=0.8)
    v = Volume(np.ones(shape, np.float32), voxwid=w, pos_c=pc)
    m = Volume(mult, voxwid=w, pos_c=pc)
    ps = cbfp.ProjectionSet(pg, data=to_gpu(y))
    cbfp.backproject_matched(cfg, v, ps, alpha=-1.5, summand_multiplier=m, block=True)
    want = orc.backproject_matched(cfg.contents, pg.to_proj_spec_array(), chunk_for(shape, w, pc),
                                   np.ones(shape), y, mult=mult, alpha=-1.5)
    err = rel_l2(v.as_numpy() - 1.0, want - 1.0)
    print(f"matched adjoint rel L2: {err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("pitch, ss", [(0.9, 1), (0.1, 2)], ids=["coarse", "fine"])
def test_column_adjoint_coarse_and_fine_detectors(device, pitch, ss):
    # upright views (the column kernel).  A coarse detector puts a few rays
    # through each voxel per view; a detector ten times finer than the
    # voxels' shadow, sub-sampled 2x2, puts hundreds through each voxel (the
    # fixed-point scale must hold many terms per voxel and view).
    rng = np.random.default_rng(17)
    n, w = 24, 0.5
    nd = 16 if pitch > 0.5 else 128
    pg = ph.circular_geometry(10, nd, nd, pitch, R=20.0, L=40.0)
    y = rng.uniform(-1, 2, (10, nd, nd)).astype(np.float32)
    cfg = cbfp.ConfigProject(ssu=ss, ssv=ss)
    v = Volume(np.zeros((n, n, n), np.float32), voxwid=w)
    ps = cbfp.ProjectionSet(pg, data=to_gpu(y))
    cbfp.backproject_matched(cfg, v, ps, block=True)
    want = orc.backproject_matched(cfg.contents, pg.to_proj_spec_array(), chunk_for((n, n, n), w),
                                   np.zeros((n, n, n)), y)
    err = rel_l2(v.as_numpy(), want)
    print(f"column adjoint pitch {pitch} ss {ss}: rel L2 {err:.3e}")
    assert np.abs(want).max() > 0
    assert err <= 1e-5


def test_tilted_adjoint_is_deterministic_and_propagates_nan(device):
    # tilted views (the general kernel) with a detector finer than the voxels'
    # shadow: many float adds per voxel, whose order varies between runs,
    # must give bit-identical 64-bit fixed-point sums.  A NaN pixel turns
    # every voxel its ray meets into NaN.
    import torch

    rng = np.random.default_rng(23)
    shape, w = (26, 22, 20), 0.5
    pg = random_geometry(rng, 7, 90, 84, R=30.0, L=60.0, pitch=0.12)
    y = rng.uniform(-1, 2, (7, 90, 84)).astype(np.float32)
    cfg = cbfp.ConfigProject(ssu=2, ssv=1)
    outs = []
    for _ in range(3):
        v = Volume(np.zeros(shape, np.float32), voxwid=w)
        cbfp.backproject_matched(cfg, v, cbfp.ProjectionSet(pg, data=to_gpu(y)), block=True)
        outs.append(v.data.clone())
    assert all(torch.equal(outs[0], o) for o in outs[1:])
    want = orc.backproject_matched(cfg.contents, pg.to_proj_spec_array(), chunk_for(shape, w), np.zeros(shape), y)
    err = rel_l2(outs[0].cpu().numpy(), want)
    print(f"tilted adjoint rel L2: {err:.3e}")
    assert np.abs(want).max() > 0 and err <= 1e-5

    y[3, 45, 40] = np.nan
    v = Volume(np.zeros(shape, np.float32), voxwid=w)
    cbfp.backproject_matched(cfg, v, cbfp.ProjectionSet(pg, data=to_gpu(y)), block=True)
    got = v.as_numpy()
    want = orc.backproject_matched(cfg.contents, pg.to_proj_spec_array(), chunk_for(shape, w), np.zeros(shape), y)
    hit = np.isnan(want)
    assert hit.any() and np.isnan(got[hit]).all()


@pytest.mark.parametrize("kind", ["upright_one_sign", "upright_signed", "tilted_signed"])
def test_adjoint_wide_dynamic_range(device, kind):
    # The detector rows below the centre are four decades dimmer than the rows
    # above, so the lower half of the volume receives only terms near 1e-4 of
    # the largest (in the same x-y columns, and for the brick kernel in the
    # same bricks, as the upper half).  Each voxel's error is measured against
    # the size of its own terms, (A^T |y|)[i]: an accumulation whose unit is
    # set by the brightest terms nearby (32-bit fixed point at the scale of a
    # bright brick) loses about 1e-5 of a dim voxel there, and a biased
    # rounding shows as a systematic error; FP32 terms summed exactly stay
    # near 1e-7.  (Relative to |A^T y| itself the error is unbounded where
    # signed terms cancel, for any FP32 weights.)
    rng = np.random.default_rng(77)
    n, w, nd, V = 32, 0.5, 48, 32
    if kind.startswith("upright"):
        pg = ph.circular_geometry(V, nd, nd, 1.0, R=40.0, L=80.0)
    else:
        pg = ph.tilted_geometry(V, nd, 1.0, inplane_deg=3.0, outplane_deg=-2.0, offset_uv=(1.5, -1.0),
                                R=40.0, L=80.0)
    y = rng.uniform(0.5, 1.0, (V, nd, nd))
    y[:, :, nd // 2:] *= 1e-4
    if kind.endswith("signed"):
        y *= rng.choice([-1.0, 1.0], size=y.shape)
    y = y.astype(np.float32)
    cfg = cbfp.ConfigProject()
    v = Volume(np.zeros((n, n, n), np.float32), voxwid=w)
    cbfp.backproject_matched(cfg, v, cbfp.ProjectionSet(pg, data=to_gpu(y)), block=True)
    got = v.as_numpy().astype(np.float64)
    specs, ch = pg.to_proj_spec_array(), chunk_for((n, n, n), w)
    want = orc.backproject_matched(cfg.contents, specs, ch, np.zeros((n, n, n)), y)
    size = orc.backproject_matched(cfg.contents, specs, ch, np.zeros((n, n, n)), np.abs(y))
    top = size.max()
    dim = (size > 1e-7 * top) & (size < 1e-3 * top)
    err_all = rel_l2(got, want)
    err_dim = float(np.linalg.norm((got - want)[dim]) / np.linalg.norm(size[dim]))
    bias = float((got[dim] - want[dim]).sum() / size[dim].sum())
    per_voxel = float(np.max(np.abs(got[dim] - want[dim]) / size[dim]))
    print(f"adjoint dynamic range ({kind}): {dim.sum()} dim voxels, rel L2 all {err_all:.3e}, dim (vs A^T|y|) "
          f"{err_dim:.3e}, relative bias {bias:.2e}, max per-voxel {per_voxel:.3e}")
    assert dim.sum() > n**3 // 8
    assert err_all <= 1e-5
    assert err_dim <= 1e-6
    assert abs(bias) <= 2e-7
    assert per_voxel <= 1e-5


def _dot_test(pg, n, w, seed):
    import torch

    from aprict_amd.autograd import Projector

    rng = np.random.default_rng(seed)
    P = Projector(pg, vol_shape=(n, n, n), voxwid=w)
    x = torch.from_numpy(rng.uniform(0, 1, (n, n, n)).astype(np.float32)).cuda()
    y = torch.from_numpy(rng.uniform(0, 1, P.sino_shape).astype(np.float32)).cuda()
    ax = P.A(x)
    aty = P.AT(y)
    lhs = float((ax.double() * y.double()).sum())
    rhs = float((x.double() * aty.double()).sum())
    return abs(lhs - rhs) / abs(lhs), lhs


def test_dot_product_cfg1(device):
    pg, _, n, w = ph.voxel_config(1)
    rel, lhs = _dot_test(pg, n, w, 1)
    print(f"cfg1 <Ax,y> = {lhs:.6e}, relative mismatch {rel:.3e}")
    assert rel < 1e-6


def test_dot_product_cfg2_full_size(device):
    pg, _, n, w = ph.voxel_config(2)
    rel, lhs = _dot_test(pg, n, w, 2)
    print(f"cfg2 (256^3 x 360 x 384^2) <Ax,y> = {lhs:.6e}, relative mismatch {rel:.3e}")
    assert rel < 1e-6


BP_CASES = {
    "exact": dict(fast_approx=False),
    "exact_zpow_soft_ss": dict(fast_approx=False, Z_on_L_factor_power=1, soften_det_edges_proportion=0.1,
                               ssx=2, ssy=1, ssz=2),
    "fast_bilinear": dict(fast_approx=True, approx_mode_bilinear_interpolate=True),
    "fast_nearest_mask_soft": dict(fast_approx=True, cyl_mask_rad=7.0, soften_det_edges_proportion=0.05),
    "fdk_direct_zval": dict(fast_approx=True, approx_mode_direct_raycast=True, Z_on_val_factor_power=-2,
                            Z_on_val_factor_val=40.0, approx_mode_bilinear_interpolate=True),
    "fast_ss": dict(fast_approx=True, approx_mode_bilinear_interpolate=True, ssx=2, ssy=2, ssz=1),
}


@pytest.mark.parametrize("name", list(BP_CASES))
def test_reference_backprojector_matches_oracle(device, name):
    rng = np.random.default_rng(len(name))
    shape, w = (24, 22, 20), 0.5
    pg = ph.circular_geometry(12, 48, 44, 1.0, R=40.0, L=80.0)
    y = rng.uniform(0, 1, (12, 48, 44)).astype(np.float32)
    cfg = cbfp.ConfigBackproject(**BP_CASES[name])
    v = Volume(np.zeros(shape, np.float32), voxwid=w)
    cbfp.backproject(cfg, v, cbfp.ProjectionSet(pg, data=to_gpu(y)), block=True)
    want = orc.backproject(cfg.contents, pg.to_proj_spec_array(), chunk_for(shape, w), np.zeros(shape), y)
    err = rel_l2(v.as_numpy(), want)
    print(f"backproject[{name}] rel L2: {err:.3e}")
    assert np.abs(want).max() > 0
    assert err <= (1e-5 if name.startswith("exact") else 2e-5)


@pytest.mark.parametrize("layout", ["v_contiguous", "u_contiguous", "fp16"])
def test_fast_backprojector_detector_edges(device, layout):
    # a volume wider and taller than the field of view: many voxels project
    # onto the detector's first and last rows and columns or just beyond
    # them, where bilinear taps fall off the detector.  v-contiguous data
    # reads the two rows of a tap column in one load; other layouts read the
    # four taps one by one; fp16 data reads pairs of halves
    import torch

    rng = np.random.default_rng(31)
    shape, w = (30, 30, 36), 0.6
    pg = ph.circular_geometry(9, 20, 16, 1.0, R=40.0, L=80.0)
    y = rng.uniform(-1, 1, (9, 20, 16)).astype(np.float32)
    cfg = cbfp.ConfigBackproject(fast_approx=True, approx_mode_bilinear_interpolate=True)
    if layout == "u_contiguous":
        data = to_gpu(y).transpose(1, 2).contiguous().transpose(1, 2)   # strides (nu nv, 1, nu)
        assert data.stride(2) != 1
        yy = y
    elif layout == "fp16":
        data = to_gpu(y).to(torch.float16)
        yy = data.float().cpu().numpy()
    else:
        data = to_gpu(y)
        yy = y
    v = Volume(np.zeros(shape, np.float32), voxwid=w)
    cbfp.backproject(cfg, v, cbfp.ProjectionSet(pg, data=data), block=True)
    want = orc.backproject(cfg.contents, pg.to_proj_spec_array(), chunk_for(shape, w), np.zeros(shape), yy)
    err = rel_l2(v.as_numpy(), want)
    print(f"fast backprojector edges ({layout}) rel L2 {err:.3e}")
    assert (want == 0).any() and np.abs(want).max() > 0   # the field of view ends inside the volume
    assert err <= 2e-5


TILTED_CASES = ["fast_bilinear", "fast_nearest_mask_soft", "fdk_direct_zval", "fast_ss"]


@pytest.mark.parametrize("layout", ["f32", "fp16"])
@pytest.mark.parametrize("name", TILTED_CASES)
def test_fast_backprojector_tilted_offset_panels(device, name, layout):
    # the fast branch's general projective map (CB_backproject.cu:478-691,
    # geometry.py:75-107) on panels tilted in and out of their plane and
    # shifted along u and v (the batched-solve geometry).  The volume is
    # wider than the field of view on one side, so some waves have only a
    # few lanes whose taps leave the detector (both branches of the
    # edge-free test run) and others have none.
    import torch

    rng = np.random.default_rng(100 + len(name))
    shape, w = (30, 26, 22), 0.6
    pg = ph.tilted_geometry(11, 44, 1.0, inplane_deg=4.0, outplane_deg=-3.0, offset_uv=(3.5, -2.5),
                            R=40.0, L=80.0)
    y = rng.uniform(0, 1, (11, 44, 44)).astype(np.float32)
    data = to_gpu(y)
    if layout == "fp16":
        data = data.to(torch.float16)
        y = data.float().cpu().numpy()
    cfg = cbfp.ConfigBackproject(**BP_CASES[name])
    v = Volume(np.zeros(shape, np.float32), voxwid=w)
    cbfp.backproject(cfg, v, cbfp.ProjectionSet(pg, data=data), block=True)
    want = orc.backproject(cfg.contents, pg.to_proj_spec_array(), chunk_for(shape, w), np.zeros(shape), y)
    err = rel_l2(v.as_numpy(), want)
    print(f"tilted backproject[{name}, {layout}] rel L2: {err:.3e}")
    assert np.abs(want).max() > 0
    assert err <= 2e-5


def test_backprojector_multiplier_and_16bit_volumes(device):
    import torch

    rng = np.random.default_rng(9)
    shape, w = (20, 20, 16), 0.5
    pg = ph.circular_geometry(10, 40, 36, 1.0, R=40.0, L=80.0)
    y = rng.uniform(0, 1, (10, 40, 36)).astype(np.float32)
    mult = rng.uniform(0, 2, shape).astype(np.float32)
    x0 = rng.uniform(0, 1, shape).astype(np.float32)
    cfg = cbfp.ConfigBackproject(fast_approx=True, apply_summand_multiplier=True,
                                 approx_mode_bilinear_interpolate=True)
    want = orc.backproject(cfg.contents, pg.to_proj_spec_array(), chunk_for(shape, w), x0, y, mult=mult)
    for dt, tol in ((np.float32, 2e-5), (np.float16, 2e-3), ("bfloat16", 1e-2)):
        v = Volume(x0, voxwid=w, dtype=dt)
        m = Volume(mult, voxwid=w, dtype=dt)
        cbfp.backproject(cfg, v, cbfp.ProjectionSet(pg, data=to_gpu(y)), summand_multiplier=m, block=True)
        err = rel_l2(v.as_numpy(), want)
        print(f"multiplier case {dt}: rel L2 {err:.3e}")
        assert err <= tol
    del torch


TALL_CASES = {
    "fast_bilinear": BP_CASES["fast_bilinear"],
    "fast_nearest_mask_soft": dict(fast_approx=True, cyl_mask_rad=1.0, soften_det_edges_proportion=0.05),
    "fdk_direct_zval": BP_CASES["fdk_direct_zval"],
    "preset2_zpow": dict(fast_approx=True, approx_mode_direct_raycast=True, approx_mode_bilinear_interpolate=True,
                         Z_on_val_factor_power=-1, Z_on_val_factor_val=40.0, backproj_weighting_preset=2),
    "zpow_L": dict(fast_approx=True, approx_mode_bilinear_interpolate=True, Z_on_L_factor_power=2),
}


def _tall_geometry(kind: str, views: int, n: int) -> cbfp.ProjGeom:
    circ = ph.circular_geometry(views, n, n, 1.0, R=40.0, L=80.0)
    tilt = ph.tilted_geometry(views, n, 1.0, inplane_deg=4.0, outplane_deg=-3.0, offset_uv=(1.5, -2.0),
                              R=40.0, L=80.0)
    if kind == "circular":
        return circ
    if kind == "tilted":
        return tilt
    # the first 64 views (one batch) upright, the rest tilted
    u = np.broadcast_to(circ.det_uvec, tilt.det_uvec.shape).copy()
    v = np.broadcast_to(circ.det_vvec, tilt.det_vvec.shape).copy()
    d = np.broadcast_to(circ.det_pos, tilt.det_pos.shape).copy()
    u[64:], v[64:], d[64:] = tilt.det_uvec[64:], tilt.det_vvec[64:], tilt.det_pos[64:]
    return cbfp.ProjGeom(source_pos=tilt.source_pos, det_pos=d, det_uvec=u, det_vvec=v, nu=n, nv=n)


@pytest.mark.parametrize("geom", ["circular", "tilted", "mixed"])
@pytest.mark.parametrize("name", list(TALL_CASES))
def test_fast_backprojector_tall_columns(device, name, geom):
    # Volumes more than two waves tall take four voxels per thread along z.  A
    # batch of 64 views whose panels are upright (z-coherent maps) shares each
    # view's u part between a thread's voxels; any other batch evaluates each
    # voxel on its own ("mixed": one batch of each).  150 planes leave the
    # fourth voxel of every thread and part of the third outside the volume,
    # and the volume is taller than the field of view, so taps leave the
    # detector's first and last rows.
    rng = np.random.default_rng(300 + len(name) + 7 * len(geom))
    views, n = 70, 40
    shape, w = (10, 9, 150), 0.25
    pg = _tall_geometry(geom, views, n)
    y = rng.uniform(0, 1, (views, n, n)).astype(np.float32)
    cfg = cbfp.ConfigBackproject(**TALL_CASES[name])
    v = Volume(np.zeros(shape, np.float32), voxwid=w)
    cbfp.backproject(cfg, v, cbfp.ProjectionSet(pg, data=to_gpu(y)), block=True)
    want = orc.backproject(cfg.contents, pg.to_proj_spec_array(), chunk_for(shape, w), np.zeros(shape), y)
    err = rel_l2(v.as_numpy(), want)
    print(f"tall-column backproject[{name}, {geom}] rel L2: {err:.3e}")
    assert (want == 0).any() and np.abs(want).max() > 0
    assert err <= 2e-5


@pytest.mark.parametrize("layout", ["u_contiguous", "fp16"])
def test_fast_backprojector_tall_columns_layouts(device, layout):
    # the same on upright panels with detector data that is not v-contiguous
    # (no paired taps: each voxel on its own) and with fp16 pairs
    import torch

    rng = np.random.default_rng(77)
    views, n = 70, 40
    shape, w = (10, 9, 200), 0.25
    pg = _tall_geometry("circular", views, n)
    y = rng.uniform(-1, 1, (views, n, n)).astype(np.float32)
    if layout == "u_contiguous":
        data = to_gpu(y).transpose(1, 2).contiguous().transpose(1, 2)
        assert data.stride(2) != 1
        yy = y
    else:
        data = to_gpu(y).to(torch.float16)
        yy = data.float().cpu().numpy()
    cfg = cbfp.ConfigBackproject(fast_approx=True, approx_mode_bilinear_interpolate=True)
    v = Volume(np.zeros(shape, np.float32), voxwid=w)
    cbfp.backproject(cfg, v, cbfp.ProjectionSet(pg, data=data), block=True)
    want = orc.backproject(cfg.contents, pg.to_proj_spec_array(), chunk_for(shape, w), np.zeros(shape), yy)
    err = rel_l2(v.as_numpy(), want)
    print(f"tall-column backproject ({layout}) rel L2 {err:.3e}")
    assert np.abs(want).max() > 0
    assert err <= 2e-5


@pytest.mark.parametrize("sino_dtype", ["float32", "float16"])
def test_matched_adjoint_helical_two_z_tiles(device, sino_dtype):
    # The upright-view adjoint takes tiles of 256 planes.  A 300-plane volume has
    # a full tile and a partial one (its top boundary is the volume's), and a
    # helical locus moves the source from below the volume to above it, so
    # tiles lie above, below and across the source plane from view to view
    # (rising, falling and both-sided evaluations).  Compared with the oracle
    # transpose, and through the dot-product identity.
    import torch

    rng = np.random.default_rng(91)
    views, nu, nv = 18, 40, 72
    shape, w = (12, 10, 300), 0.2           # 60 mm tall
    R = 40.0  # noqa: N806
    pg = ph.circular_geometry(views, nu, nv, 1.0, R=R, L=80.0)
    dz = np.linspace(-36.0, 36.0, views)     # the source runs from 6 mm below to 6 mm above the volume
    pg = cbfp.ProjGeom(source_pos=pg.source_pos + dz[:, None] * [0, 0, 1], det_pos=pg.det_pos + dz[:, None] * [0, 0, 1],
                       det_uvec=pg.det_uvec, det_vvec=np.broadcast_to(pg.det_vvec, pg.det_uvec.shape).copy(),
                       nu=nu, nv=nv)
    y = rng.uniform(-1, 2, (views, nu, nv)).astype(np.float32)
    data = to_gpu(y)
    if sino_dtype == "float16":
        data = data.to(torch.float16)
        y = data.float().cpu().numpy()
    cfg = cbfp.ConfigProject()
    v = Volume(np.zeros(shape, np.float32), voxwid=w)
    cbfp.backproject_matched(cfg, v, cbfp.ProjectionSet(pg, data=data), block=True)
    got = v.as_numpy()
    want = orc.backproject_matched(cfg.contents, pg.to_proj_spec_array(), chunk_for(shape, w), np.zeros(shape), y)
    err = rel_l2(got, want)
    print(f"helical matched adjoint, 300 planes ({sino_dtype}): rel L2 {err:.3e}")
    assert np.abs(want).max() > 0 and err <= 1e-5
    # <A x, y> = <x, A^T y> with the GPU forward
    x = rng.uniform(0, 1, shape).astype(np.float32)
    ps = cbfp.ProjectionSet(pg)
    cbfp.project(cfg, Volume(x, voxwid=w), ps, block=True)
    lhs = float(np.sum(ps.data.cpu().numpy().astype(np.float64) * y.astype(np.float64)))
    rhs = float(np.sum(x.astype(np.float64) * got.astype(np.float64)))
    print(f"  dot test {abs(lhs - rhs) / abs(lhs):.3e}")
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
