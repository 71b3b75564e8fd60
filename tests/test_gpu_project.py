"""GPU parity of the Siddon forward projector (ap_project_voxel) with the FP64 oracle.

Tolerance: FP32 volumes, relative L2 error <= 1e-5 against the FP64 oracle
(the accuracy target of the path); 16-bit storage is compared with the
oracle applied to the same rounded volume (<= 1e-5) and its storage error
against FP32 is stated (<= 3e-3 relative L2 for bf16, <= 5e-4 for fp16).
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


@pytest.fixture(scope="module")
def cfg1():
    pg, shapes, n, w = ph.voxel_config(1)
    vol = orc.voxelise_forbild(shapes, chunk_for((n, n, n), w), np.zeros((n, n, n)))
    return pg, vol.astype(np.float32), w


def gpu_project(pg, vol, w, cfg=None, dtype=np.float32, order="C", pos_c=(0, 0, 0), init=None, device="cuda"):
    import torch

    v = Volume(torch.from_numpy(np.ascontiguousarray(vol)), voxwid=w, pos_c=pos_c, order=order,
               dtype=dtype, device=device)
    data = None if init is None else torch.from_numpy(init.astype(np.float32)).to(device)
    ps = cbfp.ProjectionSet(pg, data=data, device=device)
    cbfp.project(cfg or cbfp.ConfigProject(), v, ps, block=True)
    return ps.data.cpu().numpy().astype(np.float64)


def test_forward_cfg1_fp32_matches_oracle(device, cfg1):
    pg, vol, w = cfg1
    got = gpu_project(pg, vol, w)
    want = orc.project_voxel(cbfp.ConfigProject().contents, pg.to_proj_spec_array(), chunk_for(vol.shape, w),
                             vol, np.zeros(got.shape))
    err = rel_l2(got, want)
    print(f"cfg1 forward fp32 rel L2 vs FP64 oracle: {err:.3e}")
    assert want.max() > 5
    assert err <= 1e-5


@pytest.mark.parametrize("svf,mba,ssu,ssv", [(1.0, -0.5, 2, 3), (1.25, 2.0, 1, 2)])
def test_forward_options_random_geometry(device, svf, mba, ssu, ssv):
    rng = np.random.default_rng(31)
    shape, w, pc = (40, 33, 27), 0.6, (0.4, -0.3, 0.2)
    pg = random_geometry(rng, 7, 45, 38, R=40.0, L=80.0, pitch=0.9)
    vol = rng.uniform(0, 1, shape).astype(np.float32)
    cfg = cbfp.ConfigProject(shrink_volume_factor=svf, multiply_before_adding=mba, ssu=ssu, ssv=ssv)
    init = rng.uniform(-1, 1, (7, 45, 38))
    got = gpu_project(pg, vol, w, cfg, pos_c=pc, init=init)
    want = orc.project_voxel(cfg.contents, pg.to_proj_spec_array(), chunk_for(shape, w, pc), vol,
                             init.astype(np.float32))
    err = rel_l2(got, want)
    print(f"random geometry svf={svf} ss={ssu}x{ssv} rel L2: {err:.3e}")
    assert err <= 1e-5


def test_forward_chunk_fortran_order_and_fused_residual(device):
    import torch

    rng = np.random.default_rng(4)
    shape, w = (30, 28, 26), 0.5
    pg = random_geometry(rng, 5, 40, 33, R=30.0, L=60.0, pitch=0.7)
    vol = rng.uniform(0, 1, shape).astype(np.float32)
    specs = pg.to_proj_spec_array()
    c = gpu_project(pg, vol, w)
    f = gpu_project(pg, vol, w, order="F")
    np.testing.assert_array_equal(c, f)   # same arithmetic, only the addresses differ
    # chunk: offsets (3,5,2), shape (20,17,21)
    from aprict_amd import _abi

    ch = chunk_for(shape, w, offsets=(3, 5, 2), cshape=(20, 17, 21))
    vt = torch.from_numpy(vol).cuda()
    out = torch.zeros(5, 40, 33, device="cuda")
    sp = torch.from_numpy(specs.view(np.uint8).copy()).cuda()
    cfg = cbfp.ConfigProject().contents
    _abi.check(_abi.lib().ap_project_voxel(_abi.struct_ptr(cfg), _abi.ptr(sp), 5, _abi.struct_ptr(ch),
                                           _abi.array3(vt), _abi.ptr(vt), _abi.array3(out), _abi.ptr(out),
                                           None, None, None))
    torch.cuda.synchronize()
    want = orc.project_voxel(cfg, specs, ch, vol, np.zeros((5, 40, 33)))
    assert rel_l2(out.cpu().numpy(), want) <= 1e-5
    # fused residual r = b - A x
    b = torch.from_numpy(rng.uniform(0, 5, (5, 40, 33)).astype(np.float32)).cuda()
    r = torch.full((5, 40, 33), 123.0, device="cuda")
    fp = cbfp.ConfigProject(multiply_before_adding=-1).contents
    ch0 = chunk_for(shape, w)
    _abi.check(_abi.lib().ap_project_voxel(_abi.struct_ptr(fp), _abi.ptr(sp), 5, _abi.struct_ptr(ch0),
                                           _abi.array3(vt), _abi.ptr(vt), _abi.array3(r), _abi.ptr(r),
                                           _abi.array3(b), _abi.ptr(b), None))
    torch.cuda.synchronize()
    np.testing.assert_allclose(r.cpu().numpy(), b.cpu().numpy() - c, rtol=1e-5, atol=1e-4)


@pytest.mark.parametrize("dtype,storage_tol", [("bfloat16", 3e-3), (np.float16, 5e-4)])
def test_forward_16bit_storage(device, cfg1, dtype, storage_tol):
    import torch

    pg, vol, w = cfg1
    tdt = torch.bfloat16 if dtype == "bfloat16" else torch.float16
    rounded = torch.from_numpy(vol).to(tdt).float().numpy()
    got = gpu_project(pg, vol, w, dtype=dtype)
    specs, ch = pg.to_proj_spec_array(), chunk_for(vol.shape, w)
    cfg = cbfp.ConfigProject().contents
    want_r = orc.project_voxel(cfg, specs, ch, rounded, np.zeros(got.shape))
    want = orc.project_voxel(cfg, specs, ch, vol, np.zeros(got.shape))
    e_arith, e_store = rel_l2(got, want_r), rel_l2(got, want)
    print(f"{dtype}: arithmetic error {e_arith:.3e}, storage error vs fp32 volume {e_store:.3e}")
    assert e_arith <= 1e-5
    assert e_store <= storage_tol


def test_fp16_sinogram_output(device, cfg1):
    import torch

    pg, vol, w = cfg1
    ps = cbfp.ProjectionSet(pg, data=torch.zeros(90, 96, 96, dtype=torch.float16, device="cuda"))
    cbfp.project(cbfp.ConfigProject(), Volume(vol, voxwid=w), ps, block=True)
    want = orc.project_voxel(cbfp.ConfigProject().contents, pg.to_proj_spec_array(), chunk_for(vol.shape, w),
                             vol, np.zeros((90, 96, 96)))
    assert rel_l2(ps.data.float().cpu().numpy(), want) <= 1e-3   # fp16 output rounding


def test_mixed_upright_and_tilted_views(device):
    """Upright views take the column-coherent kernels, tilted ones the general
    kernels; one call with both kinds must match the oracle, and the adjoint
    of the mixed operator must pass the dot-product test."""
    import torch

    from aprict_amd.autograd import Projector

    rng = np.random.default_rng(12)
    circ = ph.circular_geometry(9, 50, 44, 1.0, R=40.0, L=80.0)
    tilt = random_geometry(rng, 5, 50, 44, R=40.0, L=80.0, pitch=1.0)
    pg = cbfp.ProjGeom(source_pos=np.concatenate([circ.source_pos, tilt.source_pos]),
                       det_pos=np.concatenate([circ.det_pos, tilt.det_pos]),
                       det_uvec=np.concatenate([circ.det_uvec, tilt.det_uvec]),
                       det_vvec=np.concatenate([np.broadcast_to(circ.det_vvec, (9, 3)), tilt.det_vvec]),
                       nu=50, nv=44)
    shape, w = (36, 32, 30), 0.55
    vol = rng.uniform(0, 1, shape).astype(np.float32)
    for ssu, ssv in ((1, 1), (2, 3)):
        cfg = cbfp.ConfigProject(ssu=ssu, ssv=ssv)
        got = gpu_project(pg, vol, w, cfg)
        want = orc.project_voxel(cfg.contents, pg.to_proj_spec_array(), chunk_for(shape, w), vol,
                                 np.zeros(got.shape))
        err = rel_l2(got, want)
        print(f"mixed geometry ss={ssu}x{ssv} rel L2 {err:.3e}")
        assert err <= 1e-5
        P = Projector(pg, vol_shape=shape, voxwid=w, config=cfg)
        x = torch.from_numpy(vol).cuda()
        y = torch.from_numpy(rng.uniform(0, 1, P.sino_shape).astype(np.float32)).cuda()
        lhs = float((P.A(x).double() * y.double()).sum())
        rhs = float((x.double() * P.AT(y).double()).sum())
        assert abs(lhs - rhs) / abs(lhs) < 1e-6
        aty = P.AT(y).cpu().numpy()
        want_t = orc.backproject_matched(cfg.contents, pg.to_proj_spec_array(), chunk_for(shape, w),
                                         np.zeros(shape), y.cpu().numpy())
        err_t = rel_l2(aty, want_t)
        print(f"mixed geometry adjoint ss={ssu}x{ssv} rel L2 {err_t:.3e}")
        assert err_t <= 1e-5
