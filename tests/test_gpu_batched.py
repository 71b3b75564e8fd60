"""GPU checks of the launch-shape variants, the autotuner and the batched
SIRT driver (one HIP graph for many uneven problems).

- Every variant of a tunable operation gives bit-identical results (the
  decomposition changes, the arithmetic per ray / voxel does not).
- The batched graph gives bit-identical volumes to running the same
  problems eagerly, over several replays with eager work and host
  synchronisation between them (graphs have given wrong results on replays
  after the first when they held memset nodes; none are recorded here).
- The tilted fp16 problems also match the FP64 oracle's forward projection.
"""

from __future__ import annotations

import numpy as np
import pytest

from aprict_amd import _abi
from aprict_amd import phantoms as ph
from aprict_amd.autotune import Autotuner, set_variant, variant_count
from aprict_amd.cb import flatpanel as cbfp
from aprict_amd.iterators import BatchedSIRT
from aprict_amd.volume import Volume
from helpers import chunk_for, rel_l2
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _small_problems(count=5, seed=11):
    return ph.batched_problems(count, seed, n_range=(20, 44), views_range=(16, 40), det_range=(32, 72))


def _reset_variants():
    for op in range(_abi.TUNE_NUM_OPS):
        set_variant(op, 0)


def test_variants_are_bit_identical(device):
    import torch

    rng = np.random.default_rng(3)
    tilted = _small_problems(1, 21)[0]["geometry"]
    upright = ph.circular_geometry(12, 40, 150, 1.0, R=40.0, L=80.0)
    shape, w = (30, 28, 26), 0.6
    vol = Volume(torch.from_numpy(rng.uniform(0, 1, shape).astype(np.float32)), voxwid=w)
    try:
        for pg, op in ((tilted, _abi.TUNE_PROJECT_RAY), (upright, _abi.TUNE_PROJECT_COLUMN)):
            outs = []
            for var in range(variant_count(op)):
                set_variant(op, var)
                ps = cbfp.ProjectionSet(pg)
                cbfp.project(cbfp.ConfigProject(ssu=2, ssv=1), vol, ps, block=True)
                outs.append(ps.data.cpu().numpy())
            for o in outs[1:]:
                np.testing.assert_array_equal(o, outs[0])
            assert np.abs(outs[0]).max() > 0
        ps = cbfp.ProjectionSet(tilted)
        cbfp.project(cbfp.ConfigProject(), vol, ps, block=True)
        outs = []
        for var in range(variant_count(_abi.TUNE_BACKPROJECT)):
            set_variant(_abi.TUNE_BACKPROJECT, var)
            v = vol.zeroed_copy()
            cbfp.backproject(cbfp.ConfigBackproject(fast_approx=True), v, ps, block=True)
            outs.append(v.data.cpu().numpy())
        for o in outs[1:]:
            np.testing.assert_array_equal(o, outs[0])
    finally:
        _reset_variants()


def test_tilted_fp16_problem_matches_oracle(device):
    entry = _small_problems(1, 31)[0]
    x0, meas, truth = ph.batched_problem_data(entry, "float16")
    want = orc.project_voxel(cbfp.ConfigProject().contents, entry["geometry"].to_proj_spec_array(),
                             chunk_for(truth.shape, entry["voxwid"]), truth.data.cpu().numpy(),
                             np.zeros(tuple(meas.data.shape)))
    got = meas.data.float().cpu().numpy()
    err = rel_l2(got, want)
    print(f"tilted fp16 measurements vs FP64 oracle: rel L2 {err:.3e}")
    assert err < 1e-3   # fp16 storage of the sinogram
    del x0


def _make(entries, storage):
    return [ph.batched_problem_data(e, storage)[:2] for e in entries]


def test_batched_graph_equals_eager_over_replays(device):
    import torch

    entries = _small_problems(5, 11)
    tuner = Autotuner(repeats=1)
    try:
        batched = BatchedSIRT(_make(entries, "float16"), streams=3, tuner=tuner)
        eager = BatchedSIRT(_make(entries, "float16"), streams=3, tuner=tuner)
        for s in batched.solvers + eager.solvers:
            assert s.volume.data.abs().max().item() == 0
        batched.capture(3)
        assert batched.captured_iterations == 3
        for replay in range(3):
            batched.run()
            # eager work and a host synchronisation between replays
            eager.run_eager(3)
            junk = torch.randn(1 << 20, device="cuda").sum()
            torch.cuda.synchronize()
            del junk
            for i, (a, b) in enumerate(zip(batched.volumes, eager.volumes, strict=True)):
                assert torch.equal(a.data, b.data), f"replay {replay}, problem {i} differs from eager"
        # the reconstructions moved towards the truth: the residual shrinks
        r0 = [float(torch.linalg.vector_norm(s._meas.data.double())) for s in batched.solvers]
        r1 = [float(torch.linalg.vector_norm(s.residual.double())) for s in batched.solvers]
        assert all(b < a for a, b in zip(r0, r1, strict=True))
        for op in range(_abi.TUNE_NUM_OPS):
            for key in {k for keys in batched._keys for o, k in keys.items() if o == op}:
                best = tuner.best(op, key)
                assert best is not None and 0 <= best < variant_count(op)
    finally:
        _reset_variants()
