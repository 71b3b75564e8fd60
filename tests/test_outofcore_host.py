"""Out-of-core driver, host side (no GPU): brick tiling, descriptors of a
brick inside the whole volume, argument checks."""

from __future__ import annotations

import numpy as np
import pytest

from aprict_amd import outofcore as ooc
from aprict_amd import phantoms as ph


def test_bricks_tile_every_voxel_once():
    shape = (37, 20, 65)
    seen = np.zeros(shape, np.int32)
    for off, size in ooc.bricks(shape, (16, 16, 32)):
        assert all(s >= 1 for s in size)
        sl = tuple(slice(o, o + s) for o, s in zip(off, size, strict=True))
        seen[sl] += 1
    assert np.all(seen == 1)
    assert len(list(ooc.bricks(shape, (64, 64, 64)))) == 2   # z: 64 + 1
    with pytest.raises(ValueError):
        list(ooc.bricks(shape, (16, 0, 8)))


def test_shifted_descriptor_addresses_the_brick():
    torch = pytest.importorskip("torch")
    full = (50, 40, 30)
    off, size = (16, 8, 24), (16, 16, 6)
    dev = torch.zeros(size, dtype=torch.float32)           # any tensor: only addresses are used
    d, base = ooc._shifted(dev, off, full)
    assert tuple(d.shape) == full
    assert tuple(d.strides) == tuple(dev.stride())
    # element (i, j, k) of the whole volume lives at base + (i s0 + j s1 + k s2) * 4
    for i, j, k in ((16, 8, 24), (31, 23, 29), (20, 10, 25)):
        addr = base.value + (i * d.strides[0] + j * d.strides[1] + k * d.strides[2]) * 4
        li, lj, lk = i - off[0], j - off[1], k - off[2]
        assert addr == dev[li, lj, lk].data_ptr()


def test_host_objects_check_their_arguments(tmp_path):
    with pytest.raises(ValueError):
        ooc.HostVolume(np.zeros((4, 4), np.float32))
    with pytest.raises(ValueError):
        ooc.HostVolume(np.zeros((4, 4, 4), np.float64))
    with pytest.raises(ValueError):
        ooc.HostVolume(np.zeros((4, 4, 4), np.float32), voxwid=0.0)
    v = ooc.HostVolume.zeros((6, 5, 4), 0.5, (1.0, 0.0, -1.0), path=tmp_path / "v.npy")
    assert isinstance(v.data, np.memmap) and v.shape == (6, 5, 4)
    c = v.chunk_spec((2, 0, 1), (3, 5, 2))
    assert tuple(c["offsets"].tolist()) == (2, 0, 1) and tuple(c["shape"].tolist()) == (3, 5, 2)
    # texpos: corner of voxel 000 of the whole volume
    np.testing.assert_allclose(tuple(c["texpos"].tolist()), (1.0 - 1.5, -1.25, -1.0 - 1.0))
    pg = ph.circular_geometry(5, 8, 6, 1.0, R=50.0, L=100.0)
    with pytest.raises(ValueError):
        ooc.HostProjections(pg, np.zeros((5, 6, 8), np.float32))
    ooc.HostProjections(pg, np.zeros((5, 8, 6), np.float32))
