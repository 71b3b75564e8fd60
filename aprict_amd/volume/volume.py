"""Voxel and FORBILD volumes (drop-in for ``aprict.volume``).

``Volume`` mirrors ``src/aprict/volume/volume.py:51-366``: a ``[Nx][Ny][Nz]``
grid of voxels of width ``voxwid`` centred at ``pos_c``, stored in C ("z
fastest") or Fortran order.  The reference keeps it in a CUDA array behind a
surface object; here it is a torch tensor in HBM, and element-wise operations
(fill, multiply, soft cylinder mask, norms) run as HIP kernels of
``libaprict_hip.so``.  Supported storage: float32, float16 and bfloat16.

``FBVolume`` mirrors ``volume.py:369-877``: an ordered list of FORBILD
primitives (``FORBILD_Shape`` structs, last shape wins where they overlap).
"""

from __future__ import annotations

import ctypes
import math
import pathlib
import typing

import numpy as np

from .. import _abi

_ShapeType = tuple[int, int, int]


def _torch():
    import torch  # noqa: PLC0415

    return torch


def _to_torch_dtype(dtype):
    torch = _torch()
    if isinstance(dtype, torch.dtype):
        if dtype not in (torch.float16, torch.float32, torch.bfloat16):
            msg = f"unsupported volume dtype {dtype}"
            raise ValueError(msg)
        return dtype
    if isinstance(dtype, str) and dtype in ("bfloat16", "bf16"):
        return torch.bfloat16
    dt = np.dtype(dtype)
    if dt == np.float16:
        return torch.float16
    if dt == np.float32:
        return torch.float32
    msg = f"unsupported volume dtype {dt} (float16, float32 or bfloat16)"
    raise ValueError(msg)


def _public_dtype(tdt):
    """numpy dtype where numpy has one (float16/float32), else the torch dtype."""
    torch = _torch()
    if tdt == torch.float32:
        return np.dtype(np.float32)
    if tdt == torch.float16:
        return np.dtype(np.float16)
    return tdt


class Volume:
    """A rectangular prism of voxels held in GPU memory."""

    def __init__(self, arr, *, voxwid: float, pos_c=(0.0, 0.0, 0.0), initialize_texture: bool = True,
                 order: typing.Literal["C", "F"] = "C", device=None, dtype=None) -> None:
        torch = _torch()
        src = arr if isinstance(arr, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(arr))
        if src.dim() != 3:
            msg = "A volume array must be 3-dimensional."
            raise ValueError(msg)
        self._tdtype = _to_torch_dtype(dtype if dtype is not None else src.dtype)
        self._shape = tuple(int(s) for s in src.shape)
        self._order = order
        self._voxwid = float(voxwid)
        self._pos_c = np.array(pos_c, dtype=np.float32)
        self._device = torch.device(device) if device is not None else (
            src.device if src.is_cuda else torch.device("cuda"))
        self._data = None
        if initialize_texture:
            self._data = self._alloc()
            self.copy_from(src)

    # ---- storage ----------------------------------------------------------
    def _alloc(self):
        torch = _torch()
        nx, ny, nz = self._shape
        if self._order == "C":
            return torch.empty((nx, ny, nz), dtype=self._tdtype, device=self._device)
        if self._order == "F":
            return torch.empty((nz, ny, nx), dtype=self._tdtype, device=self._device).permute(2, 1, 0)
        msg = f"order must be 'C' or 'F', got {self._order!r}"
        raise ValueError(msg)

    @property
    def data(self):
        """The ``[Nx][Ny][Nz]`` torch tensor (a strided view for order 'F')."""
        if self._data is None:
            msg = "this volume was created without storage (initialize_texture=False)"
            raise RuntimeError(msg)
        return self._data

    @property
    def texture(self):
        """Alias of :attr:`data` (the reference exposes its CUDA texture here)."""
        return self.data

    @property
    def voxwid(self) -> float:
        return self._voxwid

    @property
    def shape(self) -> _ShapeType:
        return self._shape

    @property
    def dtype(self):
        return _public_dtype(self._tdtype)

    @property
    def torch_dtype(self):
        return self._tdtype

    @property
    def order(self) -> str:
        return self._order

    @property
    def pos_c(self) -> np.ndarray:
        return self._pos_c.copy()

    @property
    def device(self):
        return self._device

    @property
    def norm(self) -> float:
        """l2 norm, accumulated in FP64 on the GPU."""
        return math.sqrt(max(self.inner_product(self), 0.0))

    def inner_product(self, other: Volume) -> float:
        torch = _torch()
        out = torch.empty(1, dtype=torch.float64, device=self._device)
        _abi.check(_abi.lib().ap_array_dot(_abi.array3(self.data), _abi.ptr(self.data),
                                           _abi.array3(other.data), _abi.ptr(other.data),
                                           _abi.ptr(out), _abi.stream_handle(self._device)), "dot")
        return float(out.item())

    # ---- geometry ---------------------------------------------------------
    @property
    def xs_ys_zs(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Voxel-centre coordinates along each axis."""
        return tuple(
            self._pos_c[a] + self._voxwid * (np.arange(n, dtype=np.float32) - (n - 1) / 2)
            for a, n in enumerate(self._shape)
        )

    @property
    def minxyz_maxxyz(self):
        half = self._voxwid * np.array(self._shape, dtype=np.float64) / 2
        lo = self._pos_c.astype(np.float64) - half
        hi = self._pos_c.astype(np.float64) + half
        return tuple(float(x) for x in lo), tuple(float(x) for x in hi)

    # ---- element-wise operations (HIP kernels) ------------------------------
    def _stream(self):
        return _abi.stream_handle(self._device)

    def fill(self, value: float) -> None:
        _abi.check(_abi.lib().ap_array_fill(_abi.array3(self.data), _abi.ptr(self.data), float(value),
                                            self._stream()), "fill")

    def multiply(self, f: float) -> None:
        _abi.check(_abi.lib().ap_array_scale(_abi.array3(self.data), _abi.ptr(self.data), float(f),
                                             self._stream()), "multiply")

    def add(self, other: Volume, alpha: float = 1.0) -> None:
        """self += alpha * other."""
        _abi.check(_abi.lib().ap_array_axpby(_abi.array3(self.data), _abi.ptr(self.data), 1.0,
                                             _abi.array3(other.data), _abi.ptr(other.data), float(alpha),
                                             self._stream()), "add")

    def multiply_volume(self, other: Volume) -> None:
        _abi.check(_abi.lib().ap_array_mul(_abi.array3(self.data), _abi.ptr(self.data),
                                           _abi.array3(other.data), _abi.ptr(other.data),
                                           self._stream()), "multiply_volume")

    def multiply_by_xygrid(self, xygrid) -> None:
        torch = _torch()
        g = xygrid if isinstance(xygrid, torch.Tensor) else torch.from_numpy(np.asarray(xygrid))
        if tuple(g.shape) != self._shape[:2]:
            msg = "Input array should be 2-dimensional and match the shape[:2] of this volume."
            raise ValueError(msg)
        g3 = g.to(device=self._device, dtype=self._tdtype)[:, :, None].expand(*self._shape)
        _abi.check(_abi.lib().ap_array_mul(_abi.array3(self.data), _abi.ptr(self.data), _abi.array3(g3),
                                           _abi.ptr(g3), self._stream()), "multiply_by_xygrid")

    def divide(self, vol: Volume) -> None:
        inv = vol.copy()
        _abi.check(_abi.lib().ap_array_invert(_abi.array3(inv.data), _abi.ptr(inv.data), 0, 0.0,
                                              self._stream()), "invert")
        self.multiply_volume(inv)

    def divide_softzero(self, vol: Volume, eps: float) -> None:
        inv = vol.copy()
        _abi.check(_abi.lib().ap_array_invert(_abi.array3(inv.data), _abi.ptr(inv.data), 0, float(eps),
                                              self._stream()), "invert")
        self.multiply_volume(inv)

    def invert(self, *, ignore_zeros: bool = False, softmax_epsilon: float | None = None) -> None:
        _abi.check(_abi.lib().ap_array_invert(_abi.array3(self.data), _abi.ptr(self.data),
                                              int(ignore_zeros), float(softmax_epsilon or 0.0),
                                              self._stream()), "invert")

    def soften_to_xy_cylinder(self, relrad_start: float) -> None:
        """Raised-cosine mask to the x/y-inscribed cylinder from ``relrad_start``."""
        if not 0 <= relrad_start <= 1:
            msg = "The input 'relrad_start' should be between 0 and 1."
            raise ValueError(msg)
        _abi.check(_abi.lib().ap_volume_softmask_cylinder(_abi.array3(self.data), _abi.ptr(self.data),
                                                          float(relrad_start), self._stream()),
                   "softmask")

    # ---- copies (torch is the carrier) --------------------------------------
    def copy_from(self, arr, block: bool = True) -> None:
        torch = _torch()
        src = arr if isinstance(arr, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(arr))
        if tuple(src.shape) != self._shape:
            msg = f"shape mismatch: {tuple(src.shape)} vs {self._shape}"
            raise ValueError(msg)
        self.data.copy_(src.to(device=self._device, dtype=self._tdtype))
        if block and self._device.type == "cuda":
            torch.cuda.synchronize(self._device)

    def copy_to(self, arr, block: bool = True) -> None:
        torch = _torch()
        if isinstance(arr, torch.Tensor):
            arr.copy_(self.data)
        else:
            arr[...] = self.as_numpy()
        if block and self._device.type == "cuda":
            torch.cuda.synchronize(self._device)

    def as_numpy(self, out: np.ndarray | None = None, allow_indirect_copy: bool = False) -> np.ndarray:
        del allow_indirect_copy
        torch = _torch()
        t = self.data
        if self._tdtype == torch.bfloat16:
            t = t.float()
        x = t.cpu().numpy()
        if out is not None:
            out[...] = x
            return out
        return np.asfortranarray(x) if self._order == "F" else np.ascontiguousarray(x)

    def as_torch(self):
        return self.data.clone()

    def save(self, path) -> None:
        np.save(pathlib.Path(path), self.as_numpy())

    def sum_along_axis(self, axis: int):
        return self.data.float().sum(dim=axis)

    def copy(self) -> Volume:
        v = Volume.zeros(self._shape, voxwid=self._voxwid, dtype=self._tdtype, pos_c=self._pos_c,
                         order=self._order, device=self._device, fill=False)
        v.data.copy_(self.data)
        return v

    @classmethod
    def zeros(cls, shape: _ShapeType, *, voxwid: float, dtype=np.float32, pos_c=(0.0, 0.0, 0.0),
              initialize_texture: bool = True, order: typing.Literal["C", "F"] = "C", device=None,
              fill: bool = True) -> Volume:
        torch = _torch()
        if type(shape) is not tuple or len(shape) != 3:
            msg = "Invalid shape provided. It should be a tuple of three positive integers."
            raise ValueError(msg)
        self = cls.__new__(cls)
        self._shape = tuple(int(s) for s in shape)
        self._voxwid = float(voxwid)
        self._tdtype = _to_torch_dtype(dtype)
        self._pos_c = np.array(pos_c, dtype=np.float32)
        self._order = order
        self._device = torch.device(device) if device is not None else torch.device("cuda")
        self._data = None
        if initialize_texture:
            self._data = self._alloc()
            if fill:
                self.fill(0.0)
        return self

    def zeroed_copy(self) -> Volume:
        return Volume.zeros(self._shape, voxwid=self._voxwid, dtype=self._tdtype, pos_c=self._pos_c,
                            order=self._order, device=self._device)

    def chunk_spec(self) -> np.ndarray:
        """The whole-volume ``vol_tex_chunk_spec`` (texpos = corner of voxel 000)."""
        texpos = self._pos_c.astype(np.float64) - self._voxwid * np.array(self._shape, np.float64) / 2
        return np.array((self._voxwid, (0, 0, 0), self._shape, tuple(texpos)),
                        dtype=_abi.vol_tex_chunk_spec)


class FBVolume:
    """An ordered list of FORBILD primitives (later shapes win where they overlap)."""

    def __init__(self, shapes: np.ndarray) -> None:
        arr = np.asarray(shapes)
        if arr.dtype != _abi.forbild_shape:
            msg = "FBVolume expects an array of FORBILD_Shape structs"
            raise ValueError(msg)
        self._shapes = np.ascontiguousarray(arr.reshape(-1))
        self._dev = {}
        self._bboxes = None

    @property
    def shapes(self) -> np.ndarray:
        return self._shapes

    def __len__(self) -> int:
        return len(self._shapes)

    def shapes_device(self, device=None):
        torch = _torch()
        dev = torch.device(device) if device is not None else torch.device("cuda")
        key = str(dev)
        if key not in self._dev:
            raw = self._shapes.view(np.uint8).copy()
            self._dev[key] = torch.from_numpy(raw).to(dev)
        return self._dev[key]

    def bboxes(self, device=None) -> np.ndarray:
        """Per-shape AABBs ``[K, 2, 3]`` (clip planes ignored), computed on the GPU."""
        torch = _torch()
        if self._bboxes is None:
            k = len(self._shapes)
            if k == 0:
                self._bboxes = np.zeros((0, 2, 3), np.float32)
            else:
                dev = torch.device(device) if device is not None else torch.device("cuda")
                out = torch.empty(6 * k, dtype=torch.float32, device=dev)
                _abi.check(_abi.lib().ap_forbild_bboxes(_abi.ptr(self.shapes_device(dev)), k,
                                                        _abi.ptr(out), _abi.stream_handle(dev)),
                           "forbild_bboxes")
                self._bboxes = out.cpu().numpy().reshape(k, 2, 3)
        return self._bboxes

    def prune_to_bbox(self, min_xyz, max_xyz) -> FBVolume:
        """Keep only shapes whose AABB meets the box (order preserved)."""
        bb = self.bboxes()
        lo = np.asarray(min_xyz, np.float32)
        hi = np.asarray(max_xyz, np.float32)
        keep = np.all(bb[:, 1, :] >= lo, axis=1) & np.all(bb[:, 0, :] <= hi, axis=1)
        if np.all(keep):
            return self
        return FBVolume(self._shapes[keep])

    # ---- phantom editing (volume/volume.py:509-624) --------------------------
    def _replace(self, shapes: np.ndarray) -> None:
        self._shapes = np.ascontiguousarray(shapes)
        self._dev = {}
        self._bboxes = None

    def scale_len(self, factor: float) -> None:
        """Scale every length of the phantom (centres, sizes, clip-plane offsets) by ``factor``."""
        s = self._shapes.copy()
        for key in ("x", "y", "z", "r", "l", "dx", "dy", "dz", "r1", "r2"):
            s[key] *= factor
        for cas in ("cas1", "cas2", "cas3"):
            s[cas]["dotprod"] *= factor
        self._replace(s)

    @staticmethod
    def _axis_name(axis) -> str:
        names = {0: "x", 1: "y", 2: "z", "x": "x", "y": "y", "z": "z"}
        if axis not in names:
            msg = f"Unknown axis: {axis}"
            raise ValueError(msg)
        return names[axis]

    def flip(self, axis) -> None:
        """Mirror the phantom along one axis (0/1/2 or "x"/"y"/"z")."""
        a = self._axis_name(axis)
        s = self._shapes.copy()
        s[a] *= -1
        for key in ("axis", "a_x", "a_y", "a_z"):
            s[key][a] *= -1
        for cas in ("cas1", "cas2", "cas3"):
            s[cas]["axis"][a] *= -1
        self._replace(s)

    def translate(self, axis, amount: float) -> None:
        """Move the phantom by ``amount`` along one axis (clip planes move with it)."""
        a = self._axis_name(axis)
        s = self._shapes.copy()
        s[a] += amount
        for cas in ("cas1", "cas2", "cas3"):
            s[cas]["dotprod"] += amount * s[cas]["axis"][a]
        self._replace(s)

    def multiply_attenuation(self, amount: float) -> None:
        """Multiply every shape's density by ``amount``."""
        s = self._shapes.copy()
        s["rho"] *= amount
        self._replace(s)

    def superimpose_onto(self, other: FBVolume) -> FBVolume:
        """``other``'s shapes followed by this phantom's: this one wins where they overlap."""
        return FBVolume(np.concatenate((other.shapes, self._shapes)))

    def voxelise(self, volume: Volume, supersample: int = 1) -> Volume:
        """Add the point-sampled phantom into ``volume`` (add_FORBILD_sampling)."""
        chunk = volume.chunk_spec()
        _abi.check(_abi.lib().ap_voxelise_forbild(
            _abi.ptr(self.shapes_device(volume.device)), len(self._shapes), _abi.struct_ptr(chunk),
            _abi.array3(volume.data), _abi.ptr(volume.data), int(supersample),
            _abi.stream_handle(volume.device)), "voxelise_forbild")
        return volume


del ctypes
