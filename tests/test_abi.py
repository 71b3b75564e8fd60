"""The C-ABI library: loads without a GPU, exports every declared symbol,
has the reference's struct layouts, and rejects bad arguments with error
codes before touching the device (no compute calls here)."""

from __future__ import annotations

import ctypes
import pathlib
import re

import numpy as np
import pytest

import aprict_amd
from aprict_amd import _abi
from aprict_amd.cb import flatpanel as cbfp

HEADER = pathlib.Path(__file__).resolve().parents[1] / "include" / "aprict_hip.h"


def declared_functions() -> list[str]:
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(ap_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = aprict_amd.lib()
    names = declared_functions()
    assert len(names) == len(_abi.SIGNATURES)
    for n in names:
        assert hasattr(lib, n), n
        assert n in _abi.SIGNATURES, f"{n} has no ctypes signature"


def test_struct_layouts_match_reference_sizes():
    lib = aprict_amd.lib()
    want = {0: 48, 1: 60, 2: 40, 3: 16, 4: 16, 5: 60, 6: 20, 7: 156, 8: 40, 9: 640}
    for which, size in want.items():
        assert lib.ap_struct_size(which) == size
    assert lib.ap_struct_size(99) == -1
    # proj_chunk_spec_maxlen10 (the reference's 10-view launch argument): 10 x 60 B of
    # specs, the view count at 600, the data address at 608, byte strides at 616
    c10 = _abi.proj_chunk_spec_maxlen10
    assert c10.itemsize == 640
    assert [c10.fields[f][1] for f in ("projspecs", "numprojs", "projmemaddr", "projmemstrides")] == [0, 600, 608, 616]
    assert _abi.forbild_shape.fields["x"][1] == 68
    assert _abi.proj_spec.fields["dspec"][1] == 12
    assert _abi.config_cb_backproject.fields["ssx"][1] == 32


def _desc(shape, dtype=_abi.AP_F32):
    d = _abi.Array3()
    d.dtype = dtype
    st = [shape[1] * shape[2], shape[2], 1]
    for a in range(3):
        d.shape[a] = shape[a]
        d.strides[a] = st[a]
    return d


FAKE = ctypes.c_void_p(0x1000)   # never dereferenced: every call below fails validation first
FAKE2 = ctypes.c_void_p(0x2000)


def test_invalid_arguments_are_reported_not_launched():
    lib = aprict_amd.lib()
    cfg = cbfp.ConfigProject().contents
    chunk = np.array((0.5, (0, 0, 0), (4, 4, 4), (-1, -1, -1)), dtype=_abi.vol_tex_chunk_spec)
    vd, sd = _desc((4, 4, 4)), _desc((3, 5, 5))
    st = lib.ap_project_voxel(_abi.struct_ptr(cfg), None, 3, _abi.struct_ptr(chunk), vd, FAKE, sd, FAKE, None, None, None)
    assert st == _abi.AP_ERR_INVALID and b"specs" in lib.ap_last_error()
    # view count disagrees with the sinogram
    st = lib.ap_project_voxel(_abi.struct_ptr(cfg), FAKE, 2, _abi.struct_ptr(chunk), vd, FAKE, sd, FAKE, None, None, None)
    assert st == _abi.AP_ERR_INVALID and b"views" in lib.ap_last_error()
    # chunk larger than the volume
    big = np.array((0.5, (1, 0, 0), (4, 4, 4), (-1, -1, -1)), dtype=_abi.vol_tex_chunk_spec)
    st = lib.ap_project_voxel(_abi.struct_ptr(cfg), FAKE, 3, _abi.struct_ptr(big), vd, FAKE, sd, FAKE, None, None, None)
    assert st == _abi.AP_ERR_INVALID and b"chunk" in lib.ap_last_error()
    # bfloat16 sinograms are not supported
    st = lib.ap_project_voxel(_abi.struct_ptr(cfg), FAKE, 3, _abi.struct_ptr(chunk), vd, FAKE,
                              _desc((3, 5, 5), _abi.AP_BF16), FAKE, None, None, None)
    assert st == _abi.AP_ERR_UNSUPPORTED
    # a fused-residual b needs its descriptor, with sino's shape and dtype
    st = lib.ap_project_voxel(_abi.struct_ptr(cfg), FAKE, 3, _abi.struct_ptr(chunk), vd, FAKE, sd, FAKE, None,
                              FAKE2, None)
    assert st == _abi.AP_ERR_INVALID and b"descriptor" in lib.ap_last_error()
    st = lib.ap_project_voxel(_abi.struct_ptr(cfg), FAKE, 3, _abi.struct_ptr(chunk), vd, FAKE, sd, FAKE,
                              _desc((3, 5, 4)), FAKE2, None)
    assert st == _abi.AP_ERR_INVALID and b"shape" in lib.ap_last_error()
    st = lib.ap_project_voxel(_abi.struct_ptr(cfg), FAKE, 3, _abi.struct_ptr(chunk), vd, FAKE, sd, FAKE,
                              _desc((3, 5, 5), _abi.AP_F16), FAKE2, None)
    assert st == _abi.AP_ERR_INVALID and b"dtype" in lib.ap_last_error()
    # the SIRT residual buffer must match the measurements
    bpf = cbfp.ConfigBackproject(fast_approx=True).contents
    st = lib.ap_sirt_iterate(_abi.struct_ptr(bpf), FAKE, 3, _abi.struct_ptr(chunk), vd, FAKE, vd, FAKE, sd, FAKE,
                             _desc((3, 5, 6)), FAKE2, None)
    assert st == _abi.AP_ERR_INVALID and b"residual" in lib.ap_last_error()
    st = lib.ap_sirt_iterate(_abi.struct_ptr(bpf), FAKE, 3, _abi.struct_ptr(chunk), vd, FAKE, vd, FAKE, sd, FAKE,
                             sd, FAKE, None)
    assert st == _abi.AP_ERR_INVALID and b"alias" in lib.ap_last_error()
    # the GBC preset needs a secondary volume of the volume's shape and dtype
    bp = cbfp.ConfigBackproject(backproj_weighting_preset=7, Z_on_val_factor_val=500.0).contents
    st = lib.ap_backproject(_abi.struct_ptr(bp), FAKE, 3, _abi.struct_ptr(chunk), vd, FAKE, None, None, sd, FAKE,
                            None, None, None)
    assert st == _abi.AP_ERR_INVALID and b"secondary" in lib.ap_last_error()
    st = lib.ap_backproject(_abi.struct_ptr(bp), FAKE, 3, _abi.struct_ptr(chunk), vd, FAKE, None, None, sd, FAKE,
                            _desc((4, 4, 3)), FAKE2, None)
    assert st == _abi.AP_ERR_INVALID and b"shape" in lib.ap_last_error()
    # multiplier flag without a multiplier
    bp = cbfp.ConfigBackproject(apply_summand_multiplier=True).contents
    st = lib.ap_backproject(_abi.struct_ptr(bp), FAKE, 3, _abi.struct_ptr(chunk), vd, FAKE, None, None, sd, FAKE,
                            None, None, None)
    assert st == _abi.AP_ERR_INVALID
    # FBP rotation: odd padding, unknown mode; weight integral: too few heights
    st = lib.ap_fbp_rotate_weight(sd, FAKE, _desc((3, 6, 6)), FAKE2, None, 0.01, 0.1, 0, None)
    assert st == _abi.AP_ERR_INVALID and b"odd padding" in lib.ap_last_error()
    st = lib.ap_fbp_rotate_weight(sd, FAKE, _desc((3, 7, 7)), FAKE2, None, 0.01, 0.1, 9, None)
    assert st == _abi.AP_ERR_INVALID and b"mode" in lib.ap_last_error()
    st = lib.ap_gbc_weight_integ(FAKE, 4, FAKE2, 1, 10, _abi.struct_ptr(bp), None)
    assert st == _abi.AP_ERR_INVALID
    # ramp filter: in == out
    st = lib.ap_ramp_filter_rows(sd, FAKE, sd, FAKE, 1.0, 0, 0, None)
    assert st == _abi.AP_ERR_INVALID
    # supersample out of range
    st = lib.ap_voxelise_forbild(FAKE, 1, _abi.struct_ptr(chunk), vd, FAKE, 9, None)
    assert st == _abi.AP_ERR_INVALID
    with pytest.raises(ValueError):
        _abi.check(_abi.AP_ERR_INVALID, "x")
    with pytest.raises(NotImplementedError):
        _abi.check(_abi.AP_ERR_UNSUPPORTED, "x")


def test_python_layer_refuses_cpu_tensors():
    torch = pytest.importorskip("torch")
    from aprict_amd.volume import Volume

    v = Volume(np.zeros((4, 4, 4), np.float32), voxwid=1.0, device="cpu")
    pg = cbfp.ProjGeom.new_circular_scan(num_source_points=2, R=10.0, L_minus_R=10.0,
                                         pix_widhei_at_isocenter=(0.5, 0.5), num_pix_widhei=(3, 3))
    ps = cbfp.ProjectionSet(pg, data=torch.zeros(2, 3, 3), device="cpu")
    with pytest.raises(ValueError, match="GPU"):
        cbfp.project(cbfp.ConfigProject(), v, ps)


def test_launch_shape_variant_is_per_host_thread():
    import threading

    lib = aprict_amd.lib()
    assert lib.ap_tune_set(_abi.TUNE_PROJECT_COLUMN, 2) == 0
    seen = {}

    def other():
        seen["before"] = lib.ap_tune_get(_abi.TUNE_PROJECT_COLUMN)
        lib.ap_tune_set(_abi.TUNE_PROJECT_COLUMN, 1)
        seen["after"] = lib.ap_tune_get(_abi.TUNE_PROJECT_COLUMN)

    t = threading.Thread(target=other)
    t.start()
    t.join()
    assert seen == {"before": 0, "after": 1}
    assert lib.ap_tune_get(_abi.TUNE_PROJECT_COLUMN) == 2
    assert lib.ap_tune_set(_abi.TUNE_PROJECT_COLUMN, 0) == 0
