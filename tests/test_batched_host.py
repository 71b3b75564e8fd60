"""Host-side checks of the batched workload generator and the tuning interface
(no GPU needed: the tuning calls only read and write the library's variant
table)."""

from __future__ import annotations

import numpy as np
import pytest

from aprict_amd import _abi
from aprict_amd import phantoms as ph
from aprict_amd.autotune import size_class, variant_count


def _pixel(pg, i, X):
    S, D = pg.source_pos[i].astype(float), pg.det_pos[i].astype(float)
    U, V = pg.det_uvec[i].astype(float), pg.det_vvec[i].astype(float)
    n = np.cross(U, V)
    lam = ((D - S) @ n) / ((X - S) @ n)
    P = S + lam * (X - S) - D
    return P @ U / (U @ U), P @ V / (V @ V)


def test_batched_problems_are_within_the_stated_ranges():
    probs = ph.batched_problems(64, 5)
    assert len(probs) == 64
    for p in probs:
        assert 32 <= p["N"] <= 192 and 60 <= p["views"] <= 360 and 64 <= p["det"] <= 256
        assert all(abs(t) <= 5.0 for t in p["tilt_deg"]) and all(abs(o) <= 10.0 for o in p["offset_mm"])
        pg = p["geometry"]
        assert pg.projs_list_shape == (p["views"],)
        u, v = pg.det_uvec.astype(float), pg.det_vvec.astype(float)
        # orthogonal unit-pitch axes, tilted out of the upright position
        assert np.abs(np.sum(u * v, axis=1)).max() < 1e-5
        np.testing.assert_allclose(np.linalg.norm(u, axis=1), 1.0, rtol=1e-5)
        np.testing.assert_allclose(np.linalg.norm(v, axis=1), 1.0, rtol=1e-5)
    # same seed, same problems
    again = ph.batched_problems(4, 5)
    for a, b in zip(again, probs[:4], strict=True):
        assert a["N"] == b["N"] and a["voxwid"] == b["voxwid"]
        np.testing.assert_array_equal(a["geometry"].det_uvec, b["geometry"].det_uvec)


def test_batched_volumes_project_inside_the_detector():
    for p in ph.batched_problems(12, 7):
        pg, h, half = p["geometry"], p["N"] * p["voxwid"] / 2, p["det"] / 2
        for i in range(0, p["views"], max(1, p["views"] // 8)):
            for X in np.array(np.meshgrid([-h, h], [-h, h], [-h, h])).T.reshape(-1, 3):
                ju, kv = _pixel(pg, i, X)
                assert abs(ju) < half and abs(kv) < half


def test_size_class_buckets_to_powers_of_two():
    assert size_class(1, 2, 3, 64, 65) == (1, 2, 4, 64, 128)
    assert size_class(100, 360) == (128, 512)


def test_tuning_interface_counts_sets_and_rejects():
    lib = _abi.lib()
    for op in range(_abi.TUNE_NUM_OPS):
        n = variant_count(op)
        assert n == 3
        for v in range(n):
            assert lib.ap_tune_set(op, v) == 0
            assert lib.ap_tune_get(op) == v
        assert lib.ap_tune_set(op, n) != 0
        assert "variants" in lib.ap_last_error().decode()
        assert lib.ap_tune_set(op, -1) != 0
        assert lib.ap_tune_set(op, 0) == 0
    assert lib.ap_tune_count(_abi.TUNE_NUM_OPS) == -1
    assert lib.ap_tune_get(-1) == -1
    with pytest.raises(ValueError):
        variant_count(99)
