"""Pins the FP64 oracle of the forward projector (CPU only).

Checks that do not reuse the oracle's own algorithm: a dense system matrix
from brute-force slab lengths, closed-form chords, axis-parallel sums, the
sub-pixel block-mean invariant of aprict's supersampling test
(test_project_cb_FORBILD_supersampling.py:302-324), chunk semantics, and the
FP32 port of the reference's own DDA traversal
(CB_project_texture_chunked.cu:111-272).
"""

from __future__ import annotations

import numpy as np
import pytest

from aprict_amd import _abi
from aprict_amd import phantoms as ph
from aprict_amd.cb import flatpanel as cbfp
from helpers import chunk_for, pixel_point, random_geometry, rel_l2, source_point
from oracle import oracle as orc


def dense_system(specs, shape, chunk, svf=1.0, ssu=1, ssv=1):
    """A[(p,j,k), voxel] from slab lengths of every (sub-)ray with every voxel box.

    Uses the chunk's own (float32) voxwid / texpos, exactly as stored.
    """
    shape = tuple(shape)
    voxwid = float(chunk["voxwid"])
    corner = np.array([chunk["texpos"][c] for c in "xyz"], np.float64)
    svf = float(np.float32(svf))
    specs = specs.reshape(-1)
    nu, nv = int(specs[0]["dspec"]["nu"]), int(specs[0]["dspec"]["nv"])
    rows = []
    for sp in specs:
        S = source_point(sp) * svf
        for j in range(nu):
            for k in range(nv):
                row = np.zeros(shape)
                for su in range(ssu):
                    for sv in range(ssv):
                        Q = pixel_point(sp, j - 0.5 + (su + 0.5) / ssu, k - 0.5 + (sv + 0.5) / ssv) * svf
                        for idx in np.ndindex(shape):
                            lo = corner + voxwid * np.array(idx)
                            row[idx] += orc.slab_length(S, Q, lo, lo + voxwid)
                rows.append(row.ravel() / (ssu * ssv))
    return np.array(rows)


@pytest.mark.parametrize("svf,ssu,ssv", [(1.0, 1, 1), (1.3, 2, 3)])
def test_siddon_matches_dense_slab_matrix(svf, ssu, ssv):
    rng = np.random.default_rng(11)
    shape, w, pc = (5, 4, 6), 0.7, (0.3, -0.2, 0.1)
    pg = random_geometry(rng, 4, 5, 4, R=8.0, L=16.0, pitch=0.9)
    specs = pg.to_proj_spec_array()
    x = rng.uniform(0, 1, shape)
    cfg = cbfp.ConfigProject(shrink_volume_factor=svf, ssu=ssu, ssv=ssv).contents
    ch = chunk_for(shape, w, pc)
    got = orc.project_voxel(cfg, specs, ch, x, np.zeros((4, 5, 4)))
    A = dense_system(specs, shape, ch, svf, ssu, ssv)
    want = (A @ x.ravel()).reshape(4, 5, 4)
    assert np.abs(want).max() > 1.0  # rays really cross the volume
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-11)


def test_constant_volume_gives_chord_length():
    rng = np.random.default_rng(3)
    shape, w = (7, 6, 5), 0.5
    pg = random_geometry(rng, 3, 6, 6, R=6.0, L=12.0, pitch=0.6)
    specs = pg.to_proj_spec_array()
    got = orc.project_voxel(cbfp.ConfigProject().contents, specs, chunk_for(shape, w), np.full(shape, 2.5),
                            np.zeros((3, 6, 6)))
    half = w * np.array(shape) / 2
    for p, sp in enumerate(specs):
        for j in range(6):
            for k in range(6):
                chord = orc.slab_length(source_point(sp), pixel_point(sp, j, k), -half, half)
                assert got[p, j, k] == pytest.approx(2.5 * chord, rel=1e-12, abs=1e-12)


def _single_ray_geom(src, dst):
    """One view, 1x1 detector whose only pixel centre is ``dst``."""
    return cbfp.ProjGeom(source_pos=np.array([src]), det_pos=np.array([dst]),
                         det_uvec=np.array([[0.0, 1.0, 0.0]]), det_vvec=np.array([[0.0, 0.0, 1.0]]),
                         nu=1, nv=1)


def test_axis_parallel_ray_sums_a_voxel_row():
    rng = np.random.default_rng(5)
    shape, w = (6, 5, 4), 0.5
    x = rng.uniform(0, 1, shape)
    corner = -w * np.array(shape) / 2
    j, k = 3, 1
    y0, z0 = corner[1] + (j + 0.5) * w, corner[2] + (k + 0.5) * w
    specs = _single_ray_geom([-50.0, y0, z0], [50.0, y0, z0]).to_proj_spec_array()
    got = orc.project_voxel(cbfp.ConfigProject().contents, specs, chunk_for(shape, w), x, np.zeros((1, 1, 1)))
    assert got[0, 0, 0] == pytest.approx(w * x[:, j, k].sum(), rel=1e-12)


def test_ray_on_a_voxel_plane_uses_the_lower_voxel():
    # y exactly on the plane between voxel rows 1 and 2: the reference's
    # direction-0 convention (CB_project_texture_chunked.cu:189-203,244-246)
    # integrates the lower row
    shape, w = (4, 4, 4), 1.0
    x = np.zeros(shape)
    x[:, 1, 2] = 1.0
    x[:, 2, 2] = 10.0
    specs = _single_ray_geom([-50.0, 0.0, 0.5], [50.0, 0.0, 0.5]).to_proj_spec_array()
    got = orc.project_voxel(cbfp.ConfigProject().contents, specs, chunk_for(shape, w), x, np.zeros((1, 1, 1)))
    assert got[0, 0, 0] == pytest.approx(4.0, rel=1e-12)


def test_supersampling_equals_block_mean_of_finer_detector():
    rng = np.random.default_rng(7)
    shape, w = (8, 8, 8), 0.5
    x = rng.uniform(0, 1, shape)
    pg = random_geometry(rng, 3, 4, 3, R=10.0, L=20.0, pitch=1.2)
    fine = cbfp.ProjGeom(source_pos=pg.source_pos, det_pos=pg.det_pos, det_uvec=pg.det_uvec / 2,
                         det_vvec=pg.det_vvec / 2, nu=8, nv=6)
    cfg2 = cbfp.ConfigProject(ssu=2, ssv=2).contents
    coarse = orc.project_voxel(cfg2, pg.to_proj_spec_array(), chunk_for(shape, w), x, np.zeros((3, 4, 3)))
    f = orc.project_voxel(cbfp.ConfigProject().contents, fine.to_proj_spec_array(), chunk_for(shape, w), x,
                          np.zeros((3, 8, 6)))
    block = f.reshape(3, 4, 2, 3, 2).mean(axis=(2, 4))
    np.testing.assert_allclose(coarse, block, rtol=1e-6, atol=1e-9)


def test_chunk_equals_volume_zeroed_outside_it():
    rng = np.random.default_rng(9)
    shape, w = (6, 7, 5), 0.5
    x = rng.uniform(0, 1, shape)
    pg = random_geometry(rng, 3, 6, 5, R=8.0, L=16.0, pitch=0.5)
    specs = pg.to_proj_spec_array()
    off, cs = (1, 2, 0), (3, 4, 5)
    cfg = cbfp.ConfigProject().contents
    got = orc.project_voxel(cfg, specs, chunk_for(shape, w, offsets=off, cshape=cs), x, np.zeros((3, 6, 5)))
    z = np.zeros_like(x)
    z[1:4, 2:6, 0:5] = x[1:4, 2:6, 0:5]
    want = orc.project_voxel(cfg, specs, chunk_for(shape, w), z, np.zeros((3, 6, 5)))
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_accumulate_multiplier_and_fused_residual():
    rng = np.random.default_rng(2)
    shape, w = (6, 6, 6), 0.5
    x = rng.uniform(0, 1, shape)
    pg = random_geometry(rng, 2, 5, 5, R=8.0, L=16.0, pitch=0.5)
    specs = pg.to_proj_spec_array()
    base = rng.uniform(-1, 1, (2, 5, 5))
    ax = orc.project_voxel(cbfp.ConfigProject().contents, specs, chunk_for(shape, w), x, np.zeros((2, 5, 5)))
    acc = orc.project_voxel(cbfp.ConfigProject(multiply_before_adding=-0.5).contents, specs,
                            chunk_for(shape, w), x, base)
    np.testing.assert_allclose(acc, base - 0.5 * ax, rtol=1e-12, atol=1e-12)
    res = orc.project_voxel(cbfp.ConfigProject(multiply_before_adding=-1).contents, specs, chunk_for(shape, w),
                            x, np.full((2, 5, 5), 7.0), b=base)
    np.testing.assert_allclose(res, base - ax, rtol=1e-12, atol=1e-12)


def test_reference_dda_port_agrees_with_fp64_oracle():
    pg, shapes, n, w = ph.voxel_config(1)
    specs = pg.to_proj_spec_array()
    ch = chunk_for((n, n, n), w)
    vol = orc.voxelise_forbild(shapes, ch, np.zeros((n, n, n)))
    cfg = cbfp.ConfigProject().contents
    s64 = orc.project_voxel(cfg, specs, ch, vol, np.zeros((90, 96, 96)))
    s32 = orc.project_voxel_refdda(cfg, specs, ch, vol, np.zeros((90, 96, 96)))
    assert s64.max() > 5
    assert rel_l2(s32, s64) < 2e-5


def test_proj_spec_layout_matches_abi():
    pg = ph.circular_geometry(4, 6, 5, 1.0)
    s = pg.to_proj_spec_array()
    assert s.dtype == _abi.proj_spec and s.dtype.itemsize == 60
    assert s["dspec"]["nu"][0] == 6 and s["dspec"]["nv"][0] == 5
