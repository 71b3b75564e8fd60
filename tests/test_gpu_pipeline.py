"""GPU parity of the ramp filter, the FORBILD analytic projector and
voxeliser, the SIRT / Landweber solvers and the differentiable projector."""

from __future__ import annotations

import json
import pathlib

import numpy as np
import pytest

from aprict_amd import _abi
from aprict_amd import phantoms as ph
from aprict_amd.cb import flatpanel as cbfp
from aprict_amd.cb import ramp_filter as rf
from aprict_amd.iterators import Landweber_flatpanel, SIRT_flatpanel
from aprict_amd.volume import FBVolume, Volume, forbild
from helpers import chunk_for, rel_l2
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

KATS = json.loads((pathlib.Path(__file__).parent / "golden" / "ramp_kats.json").read_text())["cases"]


def to_gpu(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


# ---------------------------------------------------------------- ramp filter
@pytest.mark.parametrize("case", KATS, ids=[c["id"] for c in KATS])
def test_ramp_kats_on_gpu(device, case):
    x = np.asarray(case["input"], np.float32)
    got = rf.ramp_filter_rows(to_gpu(x[None, :, None]), 1.0).cpu().numpy()[0, :, 0]
    exact = orc.ramp_filter(x.astype(np.float64)[None, :, None], 1.0)[0, :, 0]
    want = np.array([float(v) for v in case["expected"]])
    print(f"{case['id']}: GPU vs exact {np.linalg.norm(got - exact):.3e}, vs reference vector "
          f"{np.linalg.norm(got - want):.3e}")
    assert np.linalg.norm(got - exact) <= 2e-6 * np.linalg.norm(exact)
    assert np.linalg.norm(got - want) <= 2e-6 * np.linalg.norm(want)


@pytest.mark.parametrize("fdk", [False, True])
def test_ramp_filter_matches_oracle(device, fdk):
    rng = np.random.default_rng(3)
    x = rng.standard_normal((5, 150, 70)).astype(np.float32)
    p = 0.001 if fdk else 0.7
    got = rf.ramp_filter_rows(to_gpu(x), p, fdk_cosine_weight=fdk).cpu().numpy()
    want = orc.ramp_filter(x, p, fdk=fdk)
    err = rel_l2(got, want)
    print(f"ramp fdk={fdk} rel L2 {err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("kernel", [0, 1, 2, 3], ids=["ramlak", "sinc", "hilbert", "hilbert_half"])
def test_ramp_filter_wide_rows_and_every_kernel(device, kernel):
    # rows longer than any LDS tap table (the taps are staged per input chunk),
    # odd lengths and a last block of partial outputs; each kernel of
    # ramp_kernels.py against the oracle's direct convolution
    rng = np.random.default_rng(30 + kernel)
    nu = 5003 if kernel == 0 else 1037
    x = rng.standard_normal((2, nu, 40)).astype(np.float32)
    got = rf.ramp_filter_rows(to_gpu(x), 0.3, kernel=kernel).cpu().numpy()
    want = orc.ramp_filter(x, 0.3, kernel=kernel)
    err = rel_l2(got, want)
    print(f"ramp kernel {kernel}, nu {nu}: rel L2 {err:.3e}")
    assert err <= 1e-5


def test_ramp_filterer_api(device):
    rng = np.random.default_rng(4)
    x = rng.standard_normal((7, 40, 12)).astype(np.float32)
    f = rf.RampFilterer(physpixwid=0.5, proj_2d_shape=(40, 12))
    (out,), _ = f.filter_bulk_axis1_concat(to_gpu(x))
    assert rel_l2(out.cpu().numpy(), orc.ramp_filter(x, 0.5)) <= 1e-5


# ------------------------------------------------------------------- FORBILD
def test_forbild_projector_matches_oracle(device):
    shapes = ph.random_primitives(40, 30.0, seed=3)
    pg = ph.circular_geometry(24, 64, 64, 0.5 * 8, R=500.0, L=1000.0)
    specs = pg.to_proj_spec_array()
    ps = cbfp.ProjectionSet(pg)
    cbfp.project(cbfp.ConfigProject(), FBVolume(shapes), ps, block=True)
    want = orc.project_forbild(cbfp.ConfigProject().contents_fb, specs, shapes, np.zeros((24, 64, 64)))
    err = rel_l2(ps.data.cpu().numpy(), want)
    print(f"FORBILD 40 primitives rel L2 {err:.3e}, overflow count {ps.out_of_mem_count}")
    assert np.abs(want).max() > 1.0
    assert err <= 1e-5 and ps.out_of_mem_count == 0


def test_forbild_head_phantom_file_matches_oracle(device):
    # the published head phantom parsed from its .pha file: 479 primitives
    # (cones and clip planes included), deep overlaps along most rays
    from aprict_amd.volume import pha

    shapes = pha.convert_ph_file_to_shapes(pathlib.Path(pha.__file__).resolve().parents[1] / "data" / "forbild" /
                                           "head" / "HeadPhantom.pha")
    pg = ph.circular_geometry(16, 64, 64, 0.9, R=500.0, L=1000.0)
    cfg = cbfp.ConfigProject(ssu=2, ssv=2)
    ps = cbfp.ProjectionSet(pg)
    cbfp.project(cfg, FBVolume(shapes), ps, block=True)
    want = orc.project_forbild(cfg.contents_fb, pg.to_proj_spec_array(), shapes, np.zeros((16, 64, 64)))
    err = rel_l2(ps.data.cpu().numpy(), want)
    print(f"FORBILD head phantom ({len(shapes)} primitives) rel L2 {err:.3e}")
    assert np.abs(want).max() > 10.0
    assert err <= 1e-5


def test_forbild_tapered_lung_tree_matches_oracle(device):
    # the generated lung phantom with tapered tubes: 124 shapes, 60 of them
    # cones along arbitrary axes, nested outer/inner tubes (mm)
    shapes = ph.FORBILD.LungPhantom(generation=4, taper=True, fix=True).shapes
    pg = ph.circular_geometry(12, 64, 64, 8.0, R=500.0, L=1000.0)
    ps = cbfp.ProjectionSet(pg)
    cbfp.project(cbfp.ConfigProject(), FBVolume(shapes), ps, block=True)
    want = orc.project_forbild(cbfp.ConfigProject().contents_fb, pg.to_proj_spec_array(), shapes,
                               np.zeros((12, 64, 64)))
    err = rel_l2(ps.data.cpu().numpy(), want)
    print(f"tapered lung tree ({len(shapes)} primitives) rel L2 {err:.3e}")
    assert np.abs(want).max() > 10.0
    assert err <= 1e-5


def test_forbild_many_nested_overlaps(device):
    # 60 nested spheres: central rays meet all of them; the covered set stays
    # one piece while the shapes are composited from the last to the first
    shapes = forbild.stack([forbild.sphere((0.01 * i, 0, 0), 30.0 - 0.4 * i, 0.1 + 0.01 * i) for i in range(60)])
    pg = ph.circular_geometry(3, 16, 16, 4.0, R=500.0, L=1000.0)
    ps = cbfp.ProjectionSet(pg)
    cbfp.project(cbfp.ConfigProject(ssu=2), FBVolume(shapes), ps, block=True)
    want = orc.project_forbild(cbfp.ConfigProject(ssu=2).contents_fb, pg.to_proj_spec_array(), shapes,
                               np.zeros((3, 16, 16)))
    assert rel_l2(ps.data.cpu().numpy(), want) <= 1e-5


def test_forbild_many_disjoint_pieces_use_exact_fallback(device):
    # a large sphere (index 0) holding 14 small disjoint spheres on the x axis:
    # rays along x meet the small ones first in compositing order, so their
    # covered set grows to more pieces than the per-ray registers hold and the
    # recomputing sweep takes over
    small = [forbild.sphere((-19.5 + 3.0 * i, 0.0, 0.0), 1.2, 0.5 + 0.05 * i) for i in range(14)]
    shapes = forbild.stack([forbild.sphere((0.0, 0.0, 0.0), 25.0, 0.2), *small])
    pg = ph.circular_geometry(4, 24, 16, 0.5, R=500.0, L=1000.0)
    for cfg in (cbfp.ConfigProject(), cbfp.ConfigProject(ssv=2)):
        ps = cbfp.ProjectionSet(pg)
        cbfp.project(cfg, FBVolume(shapes), ps, block=True)
        want = orc.project_forbild(cfg.contents_fb, pg.to_proj_spec_array(), shapes, np.zeros((4, 24, 16)))
        err = rel_l2(ps.data.cpu().numpy(), want)
        print(f"disjoint pieces (ssv={cfg.contents['ssv']}) rel L2 {err:.3e}")
        assert np.abs(want).max() > 10.0 and err <= 1e-5


def test_voxeliser_and_bboxes(device):
    shapes = ph.random_primitives(12, 10.0, seed=5)
    n, w = 40, 0.55
    v = Volume.zeros((n, n, n), voxwid=w)
    fb = FBVolume(shapes)
    fb.voxelise(v, supersample=2)
    want = orc.voxelise_forbild(shapes, chunk_for((n, n, n), w), np.zeros((n, n, n)), ss=2)
    got = v.as_numpy()
    mismatch = np.mean(np.abs(got - want) > 1e-6)
    print(f"voxeliser mismatched voxels: {mismatch:.2e}")
    assert mismatch < 1e-3
    bb = fb.bboxes()
    xs = (np.arange(n) - (n - 1) / 2) * w
    X, Y, Z = np.meshgrid(xs, xs, xs, indexing="ij")
    for k in range(len(shapes)):
        one = orc.voxelise_forbild(shapes[k:k + 1], chunk_for((n, n, n), w), np.zeros((n, n, n)))
        inside = one != 0
        if inside.any():
            assert X[inside].min() >= bb[k, 0, 0] - 1e-4 and X[inside].max() <= bb[k, 1, 0] + 1e-4
            assert Z[inside].min() >= bb[k, 0, 2] - 1e-4 and Z[inside].max() <= bb[k, 1, 2] + 1e-4


# ------------------------------------------------------------------- solvers
def _small_problem():
    pg = ph.circular_geometry(30, 48, 40, 1.0, R=60.0, L=120.0)
    n, w = 24, 0.8
    shapes = ph.random_ellipsoids(6, n * w / 2, seed=7)
    truth = orc.voxelise_forbild(shapes, chunk_for((n, n, n), w), np.zeros((n, n, n)))
    b = orc.project_voxel(cbfp.ConfigProject().contents, pg.to_proj_spec_array(), chunk_for((n, n, n), w), truth,
                          np.zeros((30, 48, 40)))
    return pg, n, w, truth, b


def _softmask_np(a, r0):
    nx, ny = a.shape[:2]
    i = np.arange(nx)[:, None] - (nx - 1) / 2
    j = np.arange(ny)[None, :] - (ny - 1) / 2
    rr = np.sqrt(i**2 + j**2) / min(nx, ny) * 2
    m = np.ones_like(rr)
    out = rr >= r0
    if r0 >= 1:
        m[out] = 0
    else:
        t = np.minimum((rr - r0) / (1 - r0), 1)
        m[out] = np.clip((1 + np.cos(np.pi * t[out])) / 2, 0, 1)
    return a * m[:, :, None]


def _sirt_oracle(pg, n, w, b, iters, fp32):
    """SIRT of sirt_vram.py:82-149 composed from oracle operators."""
    specs = pg.to_proj_spec_array()
    ch = chunk_for((n, n, n), w)
    fp = cbfp.ConfigProject().contents
    bp = cbfp.ConfigBackproject(fast_approx=True, cyl_mask_rad=n / 2 * w, soften_det_edges_proportion=0.05).contents

    def A(x):  # noqa: N802
        return orc.project_voxel(fp, specs, ch, x, np.zeros(b.shape), fp32=fp32).astype(np.float64)

    def B(y):  # noqa: N802
        return orc.backproject(bp, specs, ch, np.zeros((n, n, n)), y, fp32=fp32).astype(np.float64)

    s = 1.0 / n
    p = A(_softmask_np(np.full((n, n, n), s), 1.0))
    mean = p.mean()
    p /= mean / s
    s /= mean / s
    M = B(p)
    nz = M != 0
    M[nz] = 1 / (M[nz] + 0.1 * np.exp(-M[nz] / 0.1))
    M = _softmask_np(M * s, 0.95)
    xs = []
    x = np.zeros((n, n, n))
    for _ in range(iters):
        x = x + M * B(b - A(x))
        xs.append(x)
    return M, xs


def test_sirt_matches_oracle_composition(device):
    # Tolerance: this small problem's detector-edge softening is steep enough
    # that FP32 arithmetic itself moves the result by ~1e-4 (measured with the
    # FP32 build of the oracle); the GPU must stay within twice that.
    pg, n, w, truth, b = _small_problem()
    vol = Volume.zeros((n, n, n), voxwid=w)
    sirt = SIRT_flatpanel(SIRT_flatpanel.Config(), vol, cbfp.ProjectionSet(pg, data=to_gpu(b)))
    M64, x64 = _sirt_oracle(pg, n, w, b, 3, fp32=False)
    M32, x32 = _sirt_oracle(pg, n, w, b, 3, fp32=True)
    tol_m = max(2 * rel_l2(M32, M64), 1e-5)
    tol_x = max(2 * rel_l2(x32[-1], x64[-1]), 1e-5)
    err_m = rel_l2(sirt.preconditioner.as_numpy(), M64)
    for _ in range(3):
        sirt.iterate()
    err_x = rel_l2(vol.as_numpy(), x64[-1])
    res = np.linalg.norm(b - orc.project_voxel(cbfp.ConfigProject().contents, pg.to_proj_spec_array(),
                                               chunk_for((n, n, n), w), x64[-1], np.zeros(b.shape)))
    print(f"SIRT preconditioner rel L2 {err_m:.3e} (tol {tol_m:.2e}), x after 3 iterations {err_x:.3e} "
          f"(tol {tol_x:.2e}), relative residual {res / np.linalg.norm(b):.3f}")
    assert err_m <= tol_m and err_x <= tol_x and res < 0.5 * np.linalg.norm(b)


def test_landweber_reduces_residual(device):
    pg, n, w, truth, b = _small_problem()
    vol = Volume.zeros((n, n, n), voxwid=w)
    lw = Landweber_flatpanel(Landweber_flatpanel.Config(fast_backproject=False, rayleigh_quotient_num_iterations=12,
                                                        bp_soft_from_detedge_relrad=None,
                                                        soft_from_cylinder_relrad=None),
                             vol, cbfp.ProjectionSet(pg, data=to_gpu(b)))
    rq = lw.rayleigh_quotients[0]
    res = []
    for _ in range(4):
        lw.iterate()
        x = vol.as_numpy()
        r = b - orc.project_voxel(cbfp.ConfigProject().contents, pg.to_proj_spec_array(), chunk_for((n, n, n), w),
                                  x, np.zeros(b.shape))
        res.append(np.linalg.norm(r))
    print(f"Landweber lambda {rq:.4e}, residuals {[f'{r:.4e}' for r in res]}")
    assert rq > 0 and all(res[i + 1] < res[i] for i in range(3))


# --------------------------------------------------------------- autograd/TV
def test_autograd_gradient_and_finite_differences(device):
    import torch

    from aprict_amd.autograd import Projector, project, tv_gradient

    pg = ph.circular_geometry(8, 24, 20, 1.0, R=40.0, L=80.0)
    n, w = 12, 0.8
    P = Projector(pg, vol_shape=(n, n, n), voxwid=w)
    rng = np.random.default_rng(0)
    b = torch.from_numpy(rng.uniform(0, 5, P.sino_shape).astype(np.float32)).cuda()
    x = torch.from_numpy(rng.uniform(0, 1, (n, n, n)).astype(np.float32)).cuda().requires_grad_(True)
    loss = ((project(x, P) - b) ** 2).sum()
    loss.backward()
    g = x.grad.double().cpu().numpy()
    want = 2 * P.AT(P.A(x.detach()) - b).double().cpu().numpy()
    assert rel_l2(g, want) <= 1e-6
    # directional finite difference in FP64 through the oracle (quadratic: central difference exact)
    d = rng.standard_normal((n, n, n))
    specs, ch = pg.to_proj_spec_array(), chunk_for((n, n, n), w)
    x0 = x.detach().double().cpu().numpy()
    bb = b.double().cpu().numpy()

    def f(v):
        r = orc.project_voxel(cbfp.ConfigProject().contents, specs, ch, v, np.zeros(bb.shape)) - bb
        return float((r * r).sum())

    h = 1e-3
    fd = (f(x0 + h * d) - f(x0 - h * d)) / (2 * h)
    assert float((g * d).sum()) == pytest.approx(fd, rel=1e-5)
    # TV gradient vs central differences of the smoothed TV
    xt = torch.from_numpy(rng.uniform(0, 1, (6, 5, 7)).astype(np.float32)).cuda()
    gt = tv_gradient(xt, 1.0, eps=0.1).double().cpu().numpy()
    xn = xt.double().cpu().numpy()

    def tv(v):
        dx = np.zeros_like(v)
        dy = np.zeros_like(v)
        dz = np.zeros_like(v)
        dx[:-1] = v[1:] - v[:-1]
        dy[:, :-1] = v[:, 1:] - v[:, :-1]
        dz[:, :, :-1] = v[:, :, 1:] - v[:, :, :-1]
        return float(np.sqrt(dx**2 + dy**2 + dz**2 + 0.01).sum())

    e = np.zeros_like(xn)
    for idx in [(0, 0, 0), (2, 3, 4), (5, 4, 6), (3, 0, 2)]:
        e[...] = 0
        e[idx] = 1e-4
        fdv = (tv(xn + e) - tv(xn - e)) / 2e-4
        assert gt[idx] == pytest.approx(fdv, rel=2e-3, abs=1e-4)


def test_abi_dot_and_elementwise(device):
    import torch

    rng = np.random.default_rng(1)
    a = rng.standard_normal((9, 8, 7))
    b = rng.standard_normal((9, 8, 7))
    ta, tb = to_gpu(a), to_gpu(b)
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    _abi.check(_abi.lib().ap_array_dot(_abi.array3(ta), _abi.ptr(ta), _abi.array3(tb), _abi.ptr(tb), _abi.ptr(out),
                                       None))
    torch.cuda.synchronize()
    assert out.item() == pytest.approx(float((a.astype(np.float32).astype(np.float64) *
                                              b.astype(np.float32).astype(np.float64)).sum()), rel=1e-12)
    v = Volume(a.astype(np.float32), voxwid=1.0)
    v.soften_to_xy_cylinder(0.5)
    np.testing.assert_allclose(v.as_numpy(), _softmask_np(a.astype(np.float32), 0.5), rtol=1e-5, atol=1e-6)
