"""Out-of-core projection: volumes (and sinograms) larger than device memory.

The volume lives in host memory -- a numpy array or ``np.memmap`` laid out
``[Nx][Ny][Nz]`` -- and is processed in bricks.  Each brick is copied to the
GPU and projected (or backprojected) with the brick's chunk descriptor: the
``vol_tex_chunk_spec`` offsets/shape that restrict a traversal to a sub-box
(``cuda/cusrc_templates/CB_project_texture_chunked.cu:126-142``,
``CB_backproject.cu:952-1027``).  The library addresses the volume as
``base + (i*s0 + j*s1 + k*s2) * itemsize`` with whole-volume indices and
touches only elements inside the chunk, so the base it is given is the brick
buffer's address minus the brick's offset: only the brick is backed by
device memory.

Projection data may stay on the device (a ``ProjectionSet``) or in host
memory (``HostProjections``, e.g. a ``.npy`` memmap as ``flatpanel.py:67-83``
describes), in which case views are processed in slices of the geometry
(``ProjGeom.sliced``, ``flatpanel.py:387-425``), the way
``scripts/project_cbfp.py:458-464`` streams them.

Brick copies run on their own stream from pinned staging buffers (two of
each kind), so the copy of brick i+1 overlaps the kernels of brick i.

Sums over bricks are exact decompositions of the in-core operators: a
forward projection is the sum of the bricks' projections (each ray's
integral split at brick faces), a backprojection writes each voxel from its
own brick.  Results agree with the in-core operators to FP32 rounding (the
crossing constants are formed relative to the chunk's centre).
"""

from __future__ import annotations

import ctypes
import itertools
from collections.abc import Iterator, Sequence

import numpy as np

from . import _abi
from .cb import flatpanel as cbfp


def _torch():
    import torch  # noqa: PLC0415

    return torch


_NP_TO_TORCH = {np.dtype(np.float32): "float32", np.dtype(np.float16): "float16"}


class HostVolume:
    """A voxel volume in host memory, with the geometry of ``Volume``:
    ``voxwid`` (voxel width) and ``pos_c`` (position of the volume centre)."""

    def __init__(self, data: np.ndarray, voxwid: float = 1.0, pos_c: Sequence[float] = (0.0, 0.0, 0.0)) -> None:
        if data.ndim != 3:
            msg = f"a volume is 3-D, got shape {data.shape}"
            raise ValueError(msg)
        if data.dtype not in _NP_TO_TORCH:
            msg = f"host volumes are float32 or float16, got {data.dtype}"
            raise ValueError(msg)
        if not voxwid > 0:
            msg = "voxwid must be positive"
            raise ValueError(msg)
        self.data = data
        self.voxwid = float(voxwid)
        self.pos_c = np.asarray(pos_c, np.float64).reshape(3)

    @property
    def shape(self) -> tuple[int, int, int]:
        return tuple(int(s) for s in self.data.shape)

    @staticmethod
    def zeros(shape: Sequence[int], voxwid: float = 1.0, pos_c: Sequence[float] = (0.0, 0.0, 0.0), *,
              dtype=np.float32, path=None) -> HostVolume:
        """Zero volume in RAM, or as a ``.npy`` memmap at ``path``."""
        shape = tuple(int(s) for s in shape)
        if path is None:
            data = np.zeros(shape, dtype)
        else:
            data = np.lib.format.open_memmap(path, mode="w+", dtype=dtype, shape=shape)
        return HostVolume(data, voxwid, pos_c)

    def chunk_spec(self, offsets: Sequence[int], size: Sequence[int]) -> np.ndarray:
        texpos = self.pos_c - self.voxwid * np.array(self.shape, np.float64) / 2
        return np.array((self.voxwid, tuple(int(o) for o in offsets), tuple(int(s) for s in size), tuple(texpos)),
                        dtype=_abi.vol_tex_chunk_spec)


class HostProjections:
    """Projection data in host memory, ``[V][nu][nv]`` (flat projection list),
    with its geometry."""

    def __init__(self, proj_geom: cbfp.ProjGeom, data: np.ndarray) -> None:
        if len(proj_geom.projs_list_shape) != 1:
            msg = "host projections need a flat projection list"
            raise NotImplementedError(msg)
        nu, nv = int(np.max(proj_geom.nu)), int(np.max(proj_geom.nv))
        want = (proj_geom.num_projs, nu, nv)
        if tuple(data.shape) != want:
            msg = f"projection data of shape {data.shape} does not match the geometry {want}"
            raise ValueError(msg)
        if data.dtype not in _NP_TO_TORCH:
            msg = f"host projections are float32 or float16, got {data.dtype}"
            raise ValueError(msg)
        self.proj_geom = proj_geom
        self.data = data


def bricks(shape: Sequence[int], brick: Sequence[int]) -> Iterator[tuple[tuple[int, int, int], tuple[int, int, int]]]:
    """(offsets, size) of the bricks tiling ``shape``, x slowest."""
    if len(brick) != 3 or any(int(b) < 1 for b in brick):
        msg = f"brick shape must be three positive sizes, got {tuple(brick)}"
        raise ValueError(msg)
    ranges = [range(0, int(n), int(b)) for n, b in zip(shape, brick, strict=True)]
    for off in itertools.product(*ranges):
        size = tuple(min(int(b), int(n) - o) for o, b, n in zip(off, brick, shape, strict=True))
        yield tuple(off), size


def _view_slices(num: int, per: int | None) -> list[slice]:
    per = num if per is None else max(1, int(per))
    return [slice(a, min(num, a + per)) for a in range(0, num, per)]


class _Stager:
    """Double-buffered host->device (and device->host) brick transfers."""

    def __init__(self, brick: Sequence[int], dtype: np.dtype, device) -> None:
        torch = _torch()
        tdt = getattr(torch, _NP_TO_TORCH[np.dtype(dtype)])
        n = int(np.prod(brick))
        self.device = device
        self.copy_stream = torch.cuda.Stream(device)
        self.pinned = [torch.empty(n, dtype=tdt, pin_memory=True) for _ in range(2)]
        self.dev = [torch.empty(n, dtype=tdt, device=device) for _ in range(2)]
        self.freed = [None, None]     # compute done with dev[i] (event)
        self.copied = [None, None]    # pinned[i] -> dev[i] done (event)
        self.k = 0

    def _views(self, i: int, size):
        cnt = int(np.prod(size))
        return self.pinned[i][:cnt].view(*size), self.dev[i][:cnt].view(*size)

    def upload(self, host: np.ndarray, off, size):
        """Copy host[off:off+size] to a device buffer; returns the device brick
        (C-contiguous) and the event the compute stream must wait for."""
        torch = _torch()
        i = self.k
        self.k ^= 1
        pin, dev = self._views(i, size)
        if self.copied[i] is not None:
            self.copied[i].synchronize()       # pinned[i] no longer read by a copy
        sl = tuple(slice(o, o + s) for o, s in zip(off, size, strict=True))
        np.copyto(pin.numpy(), host[sl])
        with torch.cuda.stream(self.copy_stream):
            if self.freed[i] is not None:
                self.copy_stream.wait_event(self.freed[i])   # dev[i] no longer read by kernels
            dev.copy_(pin, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(self.copy_stream)
        self.copied[i] = ev
        return dev, i, ev

    def zeros(self, size):
        """A zeroed device brick (on the current stream)."""
        i = self.k
        self.k ^= 1
        _, dev = self._views(i, size)
        torch = _torch()
        if self.freed[i] is not None:
            torch.cuda.current_stream(self.device).wait_event(self.freed[i])
        dev.zero_()
        return dev, i

    def release(self, i: int) -> None:
        torch = _torch()
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(self.device))
        self.freed[i] = ev


def _shifted(dev, off, full_shape) -> tuple[_abi.Array3, ctypes.c_void_p]:
    """Descriptor and base address of a whole volume of ``full_shape`` whose
    brick at ``off`` is the C-contiguous device tensor ``dev``."""
    d = _abi.array3(dev)
    shift = 0
    for a in range(3):
        d.shape[a] = int(full_shape[a])
        shift += int(off[a]) * int(dev.stride(a))
    return d, ctypes.c_void_p(int(dev.data_ptr()) - shift * dev.element_size())


def _device_projections(projs, device):
    if isinstance(projs, cbfp.ProjectionSet):
        return projs
    msg = "expected a ProjectionSet or HostProjections"
    raise TypeError(msg)


# --------------------------------------------------------------------------
# forward projection
# --------------------------------------------------------------------------
def project_streamed(config: cbfp.ConfigProject, volume: HostVolume,
                     projs: cbfp.ProjectionSet | HostProjections, *, brick: Sequence[int] = (256, 256, 256),
                     views_per_slice: int | None = None, device=None) -> None:
    """Add the Siddon projection of the host volume into ``projs`` (the
    operator of ``flatpanel.project``), brick by brick.  With
    ``HostProjections`` the views are processed ``views_per_slice`` at a time
    and added into the host array."""
    torch = _torch()
    device = torch.device("cuda") if device is None else torch.device(device)
    if isinstance(projs, HostProjections):
        V = projs.proj_geom.num_projs
        for sl in _view_slices(V, views_per_slice):
            ps = cbfp.ProjectionSet(projs.proj_geom.sliced(sl), data=None, device=device)
            ps.create_buffer(dtype=projs.data.dtype)
            project_streamed(config, volume, ps, brick=brick, device=device)
            projs.data[sl] += ps.data.cpu().numpy()
        return
    ps = _device_projections(projs, device)
    ps.create_buffer(dtype=np.float32, existing_ok=True)
    data = ps._flat_data()
    _abi.require_device(data, "projection data")
    lib = _abi.lib()
    specs = ps.specs_device()
    cfg = config.contents
    sd = _abi.array3(data)
    st = _Stager(brick, volume.data.dtype, data.device)
    for off, size in bricks(volume.shape, brick):
        dev, i, ev = st.upload(volume.data, off, size)
        cur = torch.cuda.current_stream(data.device)
        cur.wait_event(ev)
        vd, base = _shifted(dev, off, volume.shape)
        chunk = volume.chunk_spec(off, size)
        _abi.check(lib.ap_project_voxel(_abi.struct_ptr(cfg), _abi.ptr(specs), data.shape[0],
                                        _abi.struct_ptr(chunk), vd, base, sd, _abi.ptr(data), None, None,
                                        _abi.stream_handle(data.device)), "project_streamed")
        st.release(i)
    torch.cuda.synchronize(data.device)


# --------------------------------------------------------------------------
# backprojection
# --------------------------------------------------------------------------
def _backproject_streamed(call, volume: HostVolume, projs, brick, views_per_slice, device,
                          multiplier: HostVolume | None) -> None:
    torch = _torch()
    device = torch.device("cuda") if device is None else torch.device(device)
    if multiplier is not None and multiplier.shape != volume.shape:
        msg = f"Shape mismatch between volume and multiplier: {volume.shape} vs {multiplier.shape}"
        raise ValueError(msg)
    if isinstance(projs, HostProjections):
        slices = [(sl, projs.proj_geom.sliced(sl)) for sl in _view_slices(projs.proj_geom.num_projs,
                                                                            views_per_slice)]
    else:
        if projs.data is None:
            msg = "You must provide projection data to this method."
            raise ValueError(msg)
        slices = [(None, projs)]
    st = _Stager(brick, volume.data.dtype, device)
    stm = _Stager(brick, multiplier.data.dtype, device) if multiplier is not None else None
    lib = _abi.lib()
    for off, size in bricks(volume.shape, brick):
        # the brick's current values: the operator adds into them
        dev, i, ev = st.upload(volume.data, off, size)
        cur = torch.cuda.current_stream(device)
        cur.wait_event(ev)
        mdev = None
        if stm is not None:
            mdev, mi, mev = stm.upload(multiplier.data, off, size)
            cur.wait_event(mev)
        vd, base = _shifted(dev, off, volume.shape)
        md, mbase = _shifted(mdev, off, volume.shape) if mdev is not None else (None, ctypes.c_void_p(0))
        chunk = volume.chunk_spec(off, size)
        for sl, pg_or_ps in slices:
            if sl is None:
                ps = pg_or_ps
            else:
                ps = cbfp.ProjectionSet(pg_or_ps, data=torch.from_numpy(
                    np.ascontiguousarray(projs.data[sl])).to(device, non_blocking=False))
            data = ps._flat_data()
            _abi.require_device(data, "projection data")
            call(lib, ps, data, chunk, vd, base, md, mbase)
        out = dev.cpu()   # waits for the kernels of this brick
        st.release(i)
        if stm is not None:
            stm.release(mi)
        sl3 = tuple(slice(o, o + s) for o, s in zip(off, size, strict=True))
        volume.data[sl3] = out.numpy()


def backproject_matched_streamed(config: cbfp.ConfigProject, volume: HostVolume,
                                 projs: cbfp.ProjectionSet | HostProjections, *, alpha: float = 1.0,
                                 summand_multiplier: HostVolume | None = None,
                                 brick: Sequence[int] = (256, 256, 256), views_per_slice: int | None = None,
                                 device=None) -> None:
    """``volume += alpha * (multiplier *) A^T projs`` (``flatpanel.backproject_matched``), brick by brick."""
    cfg = config.contents

    def call(lib, ps, data, chunk, vd, base, md, mbase):
        ws, wsb = _abi.matched_workspace(cfg, data.shape[0], data)
        _abi.check(lib.ap_backproject_matched(
            _abi.struct_ptr(cfg), _abi.ptr(ps.specs_device()), data.shape[0], _abi.struct_ptr(chunk), vd, base,
            md, mbase, _abi.array3(data), _abi.ptr(data), float(alpha), ws, wsb, _abi.stream_handle(data.device)),
            "backproject_matched_streamed")

    _backproject_streamed(call, volume, projs, brick, views_per_slice, device, summand_multiplier)


def backproject_streamed(config: cbfp.ConfigBackproject, volume: HostVolume,
                         projs: cbfp.ProjectionSet | HostProjections, *,
                         summand_multiplier: HostVolume | None = None, brick: Sequence[int] = (256, 256, 256),
                         views_per_slice: int | None = None, device=None) -> None:
    """Reference backprojection (``flatpanel.backproject``) into a host volume, brick by brick."""
    cfg = config.contents
    if int(cfg["backproj_weighting_preset"]) == 7:
        msg = "the GBC backprojection (preset 7) writes a secondary volume; stream it with flatpanel.backproject"
        raise NotImplementedError(msg)

    def call(lib, ps, data, chunk, vd, base, md, mbase):
        _abi.check(lib.ap_backproject(_abi.struct_ptr(cfg), _abi.ptr(ps.specs_device()), data.shape[0],
                                      _abi.struct_ptr(chunk), vd, base, md, mbase, _abi.array3(data),
                                      _abi.ptr(data), None, None, _abi.stream_handle(data.device)),
                   "backproject_streamed")

    _backproject_streamed(call, volume, projs, brick, views_per_slice, device, summand_multiplier)
