"""Host-side API of the drop-in layer (no GPU): geometry, configs, JSON."""

from __future__ import annotations

import numpy as np
import pytest

from aprict_amd import _abi
from aprict_amd.cb import flatpanel as cbfp


def circ(v=8, n=(6, 5), R=100.0, LmR=150.0, pix=0.4):
    return cbfp.ProjGeom.new_circular_scan(num_source_points=v, R=R, L_minus_R=LmR,
                                           pix_widhei_at_isocenter=(pix, pix), num_pix_widhei=n)


def test_circular_scan_conventions():
    pg = circ()
    R, L = 100.0, 250.0
    assert pg.source_pos.dtype == np.float32 and pg.nu.dtype == np.uint32
    np.testing.assert_allclose(np.linalg.norm(pg.source_pos, axis=1), R, rtol=1e-6)
    np.testing.assert_allclose(pg.det_pos, pg.source_pos * np.array([1 - L / R, 1 - L / R, 1]), rtol=1e-6, atol=1e-5)
    u = pg.det_uvec
    np.testing.assert_allclose(np.linalg.norm(u, axis=1), 0.4 * L / R, rtol=1e-6)
    np.testing.assert_allclose(np.sum(u * pg.source_pos, axis=1), 0, atol=1e-4)
    np.testing.assert_allclose(pg.det_vvec.reshape(-1, 3)[0], [0, 0, -0.4 * L / R], rtol=1e-6)
    assert pg.projs_list_shape == (8,) and pg.num_projs == 8
    with pytest.raises(ValueError):
        cbfp.ProjGeom.new_circular_scan(num_source_points=2, R=1.0, L_minus_R=1.0, pix_widhei_at_isocenter=(1, 1))


def test_json_roundtrip_and_equality(tmp_path):
    pg = circ()
    f = tmp_path / "g.json"
    pg.save(f)
    with pytest.raises(FileExistsError):
        pg.save(f)
    back = cbfp.ProjGeom.load(f)
    assert back == pg and not (back != pg)
    assert back != circ(pix=0.5)


def test_derived_geometries():
    pg = circ(v=10, n=(8, 6))
    s = pg.sliced(slice(2, 5))
    assert s.num_projs == 3
    np.testing.assert_array_equal(s.source_pos, pg.source_pos[2:5])
    assert pg.sliced([0, 9]).num_projs == 2
    b = pg.binned(n=2)
    assert int(b.nu.ravel()[0]) == 4 and int(b.nv.ravel()[0]) == 3
    np.testing.assert_allclose(b.det_uvec, pg.det_uvec * 2)
    with pytest.raises(ValueError):
        pg.binned(n=5)
    r = pg.rotated(xy_rads=np.pi / 2)
    np.testing.assert_allclose(r.source_pos[0], [0, 100, 0], atol=1e-4)
    t = pg.translated(x=1, y=2, z=3)
    np.testing.assert_allclose(t.det_pos - pg.det_pos, np.broadcast_to([1, 2, 3], pg.det_pos.shape), atol=1e-4)
    f = pg.flipped(flip_x=True, flip_y=False, flip_z=False)
    np.testing.assert_allclose(f.source_pos[:, 0], -pg.source_pos[:, 0])
    lo, hi = pg.get_bbox()
    assert lo[0] < -100 and hi[0] > 100


def test_broadcast_fields_and_spec_array():
    pg = cbfp.ProjGeom(source_pos=np.array([[0.0, -10, 0], [0, -10, 1]]), det_pos=np.array([0.0, 10, 0]),
                       det_uvec=np.array([1.0, 0, 0]), det_vvec=np.array([0.0, 0, 1]), nu=4, nv=3)
    assert pg.projs_list_shape == (2,)
    specs = pg.to_proj_spec_array()
    assert specs.dtype == _abi.proj_spec and specs.shape == (2,)
    assert specs[1]["sourcepos"]["z"] == 1.0 and specs[0]["dspec"]["pos"]["y"] == 10.0
    assert specs[1]["dspec"]["nu"] == 4 and specs[1]["dspec"]["nv"] == 3


def test_config_structs():
    c = cbfp.ConfigProject(multiply_before_adding=0, ssu=None).contents
    assert c["multiply_before_adding"] == 1 and c["ssu"] == 1 and c["shrink_volume_factor"] == 1
    assert cbfp.ConfigProject(multiply_before_adding=-1).contents["multiply_before_adding"] == -1
    b = cbfp.ConfigBackproject(fast_approx=True, cyl_mask_rad=3.5, soften_det_edges_proportion=0.05).contents
    assert not b["exact_line_integ_backproj"] and b["mask_to_cylinder"] and b["mask_to_cyl_rad"] == 3.5
    assert b["soften_edge_proportion"] == np.float32(0.05) and b["ssx"] == 1
    assert cbfp.ConfigBackproject().contents["exact_line_integ_backproj"]
    assert cbfp.ConfigProject().contents_fb["overlap_addition"] == False  # noqa: E712


def test_projection_set_validation():
    pytest.importorskip("torch")
    import torch

    pg = circ(v=3, n=(4, 5))
    with pytest.raises(NotImplementedError):
        cbfp.ProjectionSet(pg, data=torch.zeros(3, 5, 4), device="cpu")
    ps = cbfp.ProjectionSet(pg, device="cpu")
    ps.create_buffer(dtype=np.float32)
    assert tuple(ps.data.shape) == (3, 4, 5)
    with pytest.raises(RuntimeError):
        ps.create_buffer(dtype=np.float32)
    ps.create_buffer(dtype=np.float32, existing_ok=True)
    z = ps.zeroed_copy()
    assert z.data.shape == ps.data.shape


def test_bench_refuses_more_than_one_gpu():
    # bench.py measures one GPU; --gpus N > 1 exits with a message before touching a device
    import pathlib
    import subprocess
    import sys

    root = pathlib.Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, check=False)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
