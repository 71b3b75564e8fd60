"""GPU checks of the out-of-core driver: volumes held in host memory and
streamed brick by brick, projections on the device or in host memory.

- forward and matched adjoint agree with the in-core operators to FP32
  rounding (relative L2 <= 1e-6; the crossing constants are formed relative
  to each brick's centre) and with the FP64 oracle (<= 1e-5);
- the reference backprojector, whose voxels see the same arithmetic whatever
  the brick, agrees with the in-core one to <= 1e-7;
- projecting into host memory view slice by view slice gives exactly the
  device-resident result.
Bricks are ragged (the volume is not a multiple of the brick) and both the
upright-detector kernels and the general ones are exercised.
"""

from __future__ import annotations

import numpy as np
import pytest

from aprict_amd import outofcore as ooc
from aprict_amd import phantoms as ph
from aprict_amd.cb import flatpanel as cbfp
from aprict_amd.volume import Volume
from helpers import chunk_for, random_geometry, rel_l2
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SHAPE, W, PC = (50, 44, 38), 0.5, (0.3, -0.2, 0.1)
BRICK = (32, 16, 32)


def _geometry(kind):
    if kind == "upright":
        return ph.circular_geometry(9, 48, 40, 0.9, R=40.0, L=80.0)
    return random_geometry(np.random.default_rng(4), 7, 45, 41, R=40.0, L=80.0, pitch=0.9)


@pytest.mark.parametrize("kind", ["upright", "tilted"])
def test_streamed_forward_matches_in_core_and_oracle(device, kind):
    import torch

    rng = np.random.default_rng(8)
    x = rng.uniform(0, 1, SHAPE).astype(np.float32)
    pg = _geometry(kind)
    cfg = cbfp.ConfigProject(ssu=2, ssv=1)
    ps = cbfp.ProjectionSet(pg)
    ooc.project_streamed(cfg, ooc.HostVolume(x, W, PC), ps, brick=BRICK)
    got = ps.data.cpu().numpy()
    ref = cbfp.ProjectionSet(pg)
    cbfp.project(cfg, Volume(torch.from_numpy(x).cuda(), voxwid=W, pos_c=PC), ref, block=True)
    want = orc.project_voxel(cfg.contents, pg.to_proj_spec_array(), chunk_for(SHAPE, W, PC), x,
                             np.zeros(got.shape))
    e_core, e_orc = rel_l2(got, ref.data.cpu().numpy()), rel_l2(got, want)
    print(f"{kind}: streamed forward vs in-core {e_core:.3e}, vs oracle {e_orc:.3e}")
    assert np.abs(want).max() > 0
    assert e_core <= 1e-6 and e_orc <= 1e-5
    # projections kept in host memory, 4 views at a time
    host = ooc.HostProjections(pg, np.zeros(got.shape, np.float32))
    ooc.project_streamed(cfg, ooc.HostVolume(x, W, PC), host, brick=BRICK, views_per_slice=4)
    np.testing.assert_array_equal(host.data, got)


@pytest.mark.parametrize("kind", ["upright", "tilted"])
def test_streamed_matched_adjoint_matches_in_core_and_oracle(device, kind, tmp_path):
    import torch

    rng = np.random.default_rng(9)
    pg = _geometry(kind)
    nu, nv = int(np.max(pg.nu)), int(np.max(pg.nv))
    y = rng.uniform(-1, 1, (pg.num_projs, nu, nv)).astype(np.float32)
    mult = rng.uniform(0.5, 1.5, SHAPE).astype(np.float32)
    cfg = cbfp.ConfigProject(ssu=1, ssv=2)
    # the volume accumulates: start from ones, in a memmap
    vol = ooc.HostVolume.zeros(SHAPE, W, PC, path=tmp_path / "x.npy")
    vol.data[...] = 1.0
    ps = cbfp.ProjectionSet(pg, data=torch.from_numpy(y).cuda())
    ooc.backproject_matched_streamed(cfg, vol, ps, alpha=-0.5, summand_multiplier=ooc.HostVolume(mult, W, PC),
                                     brick=BRICK)
    got = np.asarray(vol.data) - 1.0
    v = Volume(np.ones(SHAPE, np.float32), voxwid=W, pos_c=PC)
    cbfp.backproject_matched(cfg, v, ps, alpha=-0.5, summand_multiplier=Volume(mult, voxwid=W, pos_c=PC),
                             block=True)
    core = v.as_numpy() - 1.0
    want = orc.backproject_matched(cfg.contents, pg.to_proj_spec_array(), chunk_for(SHAPE, W, PC),
                                   np.ones(SHAPE), y, mult=mult, alpha=-0.5) - 1.0
    e_core, e_orc = rel_l2(got, core), rel_l2(got, want)
    print(f"{kind}: streamed adjoint vs in-core {e_core:.3e}, vs oracle {e_orc:.3e}")
    assert e_core <= 1e-6 and e_orc <= 1e-5
    # projections in host memory, 3 views at a time
    vol2 = ooc.HostVolume(np.ones(SHAPE, np.float32), W, PC)
    ooc.backproject_matched_streamed(cfg, vol2, ooc.HostProjections(pg, y), alpha=-0.5,
                                     summand_multiplier=ooc.HostVolume(mult, W, PC), brick=BRICK,
                                     views_per_slice=3)
    e_sl = rel_l2(vol2.data - 1.0, core)
    print(f"{kind}: view-sliced streamed adjoint vs in-core {e_sl:.3e}")
    assert e_sl <= 1e-6


def test_streamed_reference_backprojector(device):
    import torch

    rng = np.random.default_rng(10)
    pg = _geometry("upright")
    y = rng.uniform(0, 1, (pg.num_projs, 48, 40)).astype(np.float32)
    ps = cbfp.ProjectionSet(pg, data=torch.from_numpy(y).cuda())
    for fast in (True, False):
        cfg = cbfp.ConfigBackproject(fast_approx=fast, cyl_mask_rad=9.0)
        vol = ooc.HostVolume(np.zeros(SHAPE, np.float32), W, PC)
        ooc.backproject_streamed(cfg, vol, ps, brick=BRICK)
        v = Volume(np.zeros(SHAPE, np.float32), voxwid=W, pos_c=PC)
        cbfp.backproject(cfg, v, ps, block=True)
        err = rel_l2(vol.data, v.as_numpy())
        print(f"reference backprojector fast={fast}: streamed vs in-core {err:.3e}")
        assert np.abs(v.as_numpy()).max() > 0
        assert err <= 1e-7
