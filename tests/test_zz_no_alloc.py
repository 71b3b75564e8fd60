"""The allocation-free matched adjoint (ap_backproject_matched_noalloc, a
64 MiB scratch array per GPU) against ap_backproject_matched with a workspace
of its own, through the C ABI: two calls in a row, the first with views so
large that the scratch takes them three at a time, the second smaller; and
the rejection of a detector whose single view does not fit."""

from __future__ import annotations

import numpy as np
import pytest

from aprict_amd import _abi
from aprict_amd import phantoms as ph
from aprict_amd.cb import flatpanel as cbfp
from helpers import chunk_for, rel_l2
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _adjoints(pg, n, w, y, cfg):
    import torch

    lib = _abi.lib()
    c = cfg.contents
    specs = cbfp.ProjectionSet(pg, device="cuda").specs_device()
    ch = chunk_for((n, n, n), w)
    yt = torch.from_numpy(np.ascontiguousarray(y, np.float32)).cuda()
    out = []
    for mode in ("noalloc", "workspace"):
        v = torch.zeros((n, n, n), dtype=torch.float32, device="cuda")
        st = _abi.stream_handle(yt.device)
        if mode == "noalloc":
            rc = lib.ap_backproject_matched_noalloc(_abi.struct_ptr(c), _abi.ptr(specs), y.shape[0],
                                                    _abi.struct_ptr(ch), _abi.array3(v), _abi.ptr(v), None, None,
                                                    _abi.array3(yt), _abi.ptr(yt), 1.0, st)
        else:
            nb = int(lib.ap_backproject_matched_workspace_size(_abi.struct_ptr(c), y.shape[0], _abi.array3(yt)))
            ws = torch.empty(nb, dtype=torch.uint8, device="cuda")
            rc = lib.ap_backproject_matched(_abi.struct_ptr(c), _abi.ptr(specs), y.shape[0], _abi.struct_ptr(ch),
                                            _abi.array3(v), _abi.ptr(v), None, None, _abi.array3(yt), _abi.ptr(yt),
                                            1.0, _abi.ptr(ws), nb, st)
        assert rc == 0, lib.ap_last_error()
        torch.cuda.synchronize()
        out.append(v.double().cpu().numpy())
    return out


def test_noalloc_matches_workspace_entry_point(device):
    rng = np.random.default_rng(21)
    # 1) 2048 x 1000 detector: one view needs ~26 MB of prefix sums, so the
    #    64 MiB scratch takes the 8 views in several passes
    pg = ph.circular_geometry(8, 2048, 1000, 0.05, R=60.0, L=120.0)
    n, w = 32, 0.5
    y = rng.uniform(0, 1, (8, 2048, 1000)).astype(np.float32)
    cfg = cbfp.ConfigProject()
    per_view = _abi.lib().ap_backproject_matched_workspace_size(_abi.struct_ptr(cfg.contents), 1,
                                                                _abi.array3(__import__("torch").zeros((1, 2048, 1000))))
    per_pass = (64 << 20) // per_view
    assert 2 <= per_pass < 8
    got, want = _adjoints(pg, n, w, y, cfg)
    err = rel_l2(got, want)
    print(f"noalloc vs workspace, 8 views of 2048x1000 ({per_pass} per scratch pass): rel L2 {err:.3e}")
    assert np.abs(want).max() > 0 and err <= 1e-6
    # 2) a second, smaller call on the same scratch: one pass; also against the oracle
    pg2 = ph.circular_geometry(6, 40, 36, 1.0, R=60.0, L=120.0)
    n2, w2 = 20, 0.8
    y2 = rng.uniform(-1, 2, (6, 40, 36)).astype(np.float32)
    cfg2 = cbfp.ConfigProject(ssu=2, ssv=3)
    got2, want2 = _adjoints(pg2, n2, w2, y2, cfg2)
    ora = orc.backproject_matched(cfg2.contents, pg2.to_proj_spec_array(), chunk_for((n2, n2, n2), w2),
                                  np.zeros((n2, n2, n2)), y2)
    err2, err_o = rel_l2(got2, want2), rel_l2(got2, ora)
    print(f"second call, 6 views of 40x36 ss 2x3: rel L2 vs workspace {err2:.3e}, vs oracle {err_o:.3e}")
    assert err2 <= 1e-6 and err_o <= 1e-5


def test_noalloc_rejects_views_beyond_its_scratch(device):
    import torch

    lib = _abi.lib()
    cfg = cbfp.ConfigProject(ssu=4).contents
    x = torch.zeros((4, 4, 4), dtype=torch.float32, device="cuda")
    y = torch.zeros((1, 4096, 500), dtype=torch.float32, device="cuda")
    specs = torch.zeros((1, 15), dtype=torch.float32, device="cuda")   # never read: rejected before any launch
    rc = lib.ap_backproject_matched_noalloc(_abi.struct_ptr(cfg), _abi.ptr(specs), 1,
                                            _abi.struct_ptr(chunk_for((4, 4, 4), 1.0)), _abi.array3(x), _abi.ptr(x),
                                            None, None, _abi.array3(y), _abi.ptr(y), 1.0,
                                            _abi.stream_handle(x.device))
    assert rc == _abi.AP_ERR_UNSUPPORTED
    assert b"allocation-free" in lib.ap_last_error()
