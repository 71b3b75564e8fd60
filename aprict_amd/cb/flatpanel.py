"""Flat-panel cone-beam projection and backprojection (drop-in for ``aprict.cb.flatpanel``).

Entry points, with the reference interface each one mirrors
(paths relative to aprict's ``src/aprict/``):

- ``ProjGeom``          -- ``cb/flatpanel.py:269-768``
- ``ProjectionSet``     -- ``cb/flatpanel.py:771-882``
- ``ConfigProject``     -- ``cb/flatpanel.py:885-926``
- ``ConfigBackproject`` -- ``cb/flatpanel.py:929-994``
- ``project``           -- ``cb/flatpanel.py:997-1125``
- ``backproject``       -- ``cb/flatpanel.py:1128-1281``

Data live in torch tensors on the GPU (the reference uses CuPy arrays and
CUDA surfaces); every computation is a call into ``libaprict_hip.so``.  There
is no CPU fallback: with the library missing or a CPU tensor the calls raise.

Conventions kept from the reference: ``det_pos`` is the detector centre,
``det_uvec``/``det_vvec`` join adjacent pixel centres, the projection buffer
is ``[..., nu, nv]`` with ``p[i][j+1][k]`` one ``det_uvec`` from ``p[i][j][k]``,
``project``/``backproject`` add into the caller's buffer.
"""

from __future__ import annotations

import dataclasses
import json
import math
import pathlib
import typing

import numpy as np

from .. import _abi
from ..volume import FBVolume, Volume

__all__ = [
    "ConfigBackproject",
    "ConfigProject",
    "ProjGeom",
    "ProjectionSet",
    "backproject",
    "backproject_matched",
    "project",
]


def _torch():
    import torch  # noqa: PLC0415

    return torch


def _np_to_torch_dtype(dtype):
    torch = _torch()
    if isinstance(dtype, torch.dtype):
        return dtype
    dt = np.dtype(dtype)
    table = {np.dtype(np.float16): torch.float16, np.dtype(np.float32): torch.float32,
             np.dtype(np.float64): torch.float64}
    if dt not in table:
        msg = f"unsupported projection dtype {dt}"
        raise ValueError(msg)
    return table[dt]


def _check_geometry(proj_specs: np.ndarray, data) -> tuple[int, ...]:
    """Validate a proj_spec array against a data buffer; return the buffer shape.

    Same checks as cb/flatpanel.py:235-266: proj_spec dtype, flat-panel type,
    uniform detector size, at least 1-D, and the buffer shape.
    """
    if proj_specs.dtype != _abi.proj_spec:
        msg = "Invalid dtype for the input array. It should be an array of 'proj_spec' objects."
        raise ValueError(msg)
    if proj_specs.ndim < 1:
        msg = "The input proj_specs array is not at least a 1D array. It should be at least a 1D array."
        raise ValueError(msg)
    ds = proj_specs["dspec"]
    if not np.all(ds["type"] == 0):
        msg = "Only flat-panel detectors are currently supported."
        raise NotImplementedError(msg)
    nus, nvs = ds["nu"].ravel(), ds["nv"].ravel()
    if not (np.all(nus == nus[0]) and np.all(nvs == nvs[0])):
        msg = "The input detector dimensions are inconsistent between projections. This is not currently implemented."
        raise NotImplementedError(msg)
    shape = tuple(proj_specs.shape) + (int(nus[0]), int(nvs[0]))
    if data is not None and tuple(data.shape) != shape:
        msg = f"Invalid shape for the output_array. The expected shape is {shape}, but the provided output_array has shape {tuple(data.shape)}."
        raise NotImplementedError(msg)
    return shape


# ==========================================================================
# geometry
# ==========================================================================
@dataclasses.dataclass(kw_only=True, eq=False)
class ProjGeom:
    """Geometry of a set of flat-panel cone-beam projections.

    Fields are numpy arrays whose leading axes index the projection list;
    a leading axis of length 1 broadcasts over all projections.
    """

    source_pos: np.ndarray
    det_pos: np.ndarray
    det_uvec: np.ndarray
    det_vvec: np.ndarray
    nu: np.ndarray
    nv: np.ndarray

    def __init__(self, *, source_pos, det_pos, det_uvec, det_vvec, nu, nv,
                 _skip_regularisation: bool = False) -> None:
        if _skip_regularisation:
            self.source_pos, self.det_pos = source_pos, det_pos
            self.det_uvec, self.det_vvec = det_uvec, det_vvec
            self.nu, self.nv = nu, nv
            return
        vecs = [np.asarray(a, dtype=np.float64) for a in (source_pos, det_pos, det_uvec, det_vvec)]
        ints = [np.asarray(a) for a in (nu, nv)]
        for v in vecs:
            if v.ndim < 1 or v.shape[-1] != 3:
                msg = "vector fields must have a trailing axis of length 3"
                raise ValueError(msg)
        # the projection-list rank is the largest rank among the fields;
        # lower-rank fields get trailing broadcast axes of length 1
        rank = max([v.ndim - 1 for v in vecs] + [i.ndim for i in ints])
        vecs = [v.reshape(v.shape[:-1] + (1,) * (rank - (v.ndim - 1)) + (3,)) for v in vecs]
        ints = [i.reshape(i.shape + (1,) * (rank - i.ndim)) for i in ints]
        np.broadcast_shapes(*[v.shape[:-1] for v in vecs], *[i.shape for i in ints])
        self.source_pos, self.det_pos, self.det_uvec, self.det_vvec = (v.astype(np.float32) for v in vecs)
        self.nu, self.nv = (i.astype(np.uint32) for i in ints)

    # ---- shape ----------------------------------------------------------
    @property
    def projs_list_shape(self) -> tuple[int, ...]:
        """Broadcast shape of the projection list, e.g. ``(V,)``."""
        return np.broadcast_shapes(
            self.source_pos.shape[:-1], self.det_pos.shape[:-1], self.det_uvec.shape[:-1],
            self.det_vvec.shape[:-1], self.nu.shape, self.nv.shape,
        )

    @property
    def num_projs(self) -> int:
        return math.prod(self.projs_list_shape)

    def get_bbox(self, *, index_tuple: tuple[int, ...] | None = None):
        """((minx,miny,minz),(maxx,maxy,maxz)) of sources and detector corners."""
        it = () if index_tuple is None else tuple(index_tuple)
        shp = self.projs_list_shape
        lo = np.full(3, np.inf)
        hi = np.full(3, -np.inf)
        half_u = self.det_uvec * (self.nu[..., None] / 2)
        half_v = self.det_vvec * (self.nv[..., None] / 2)
        pts = [self.source_pos]
        for su in (1, -1):
            for sv in (1, -1):
                pts.append(self.det_pos + su * half_u + sv * half_v)
        for p in pts:
            p = np.broadcast_to(p, shp + (3,))[it].reshape(-1, 3)
            lo = np.minimum(lo, p.min(axis=0))
            hi = np.maximum(hi, p.max(axis=0))
        return tuple(float(x) for x in lo), tuple(float(x) for x in hi)

    # ---- derived geometries --------------------------------------------
    def sliced(self, sl) -> ProjGeom:
        """Copy restricted to a slice / index list of a flat projection list."""
        if type(sl) not in (slice, list):
            msg = f"Invalid type for slice object {sl}"
            raise ValueError(msg)
        if len(self.projs_list_shape) != 1:
            msg = "Slicing has not been implemented for non-flat projection lists."
            raise NotImplementedError(msg)
        n = self.num_projs

        def cut(a):
            if n != 1 and a.shape[0] == 1:
                return a.copy()
            return a[sl, ...].copy()

        return ProjGeom(source_pos=cut(self.source_pos), det_pos=cut(self.det_pos),
                        det_uvec=cut(self.det_uvec), det_vvec=cut(self.det_vvec),
                        nu=cut(self.nu), nv=cut(self.nv))

    def flipped(self, *, flip_x: bool, flip_y: bool, flip_z: bool) -> ProjGeom:
        f = np.array([-1 if flip_x else 1, -1 if flip_y else 1, -1 if flip_z else 1], np.float32)
        return ProjGeom(source_pos=self.source_pos * f, det_pos=self.det_pos * f,
                        det_uvec=self.det_uvec * f, det_vvec=self.det_vvec * f, nu=self.nu, nv=self.nv)

    def rotated(self, *, xy_rads: float) -> ProjGeom:
        c, s = np.cos(xy_rads), np.sin(xy_rads)
        rot = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]], dtype=np.float32)
        return ProjGeom(source_pos=self.source_pos @ rot.T, det_pos=self.det_pos @ rot.T,
                        det_uvec=self.det_uvec @ rot.T, det_vvec=self.det_vvec @ rot.T,
                        nu=self.nu, nv=self.nv)

    def translated(self, *, x: float, y: float, z: float) -> ProjGeom:
        t = np.array([x, y, z], dtype=np.float32)
        return ProjGeom(source_pos=self.source_pos + t, det_pos=self.det_pos + t,
                        det_uvec=self.det_uvec, det_vvec=self.det_vvec, nu=self.nu, nv=self.nv)

    def binned(self, *, n: int) -> ProjGeom:
        if not n > 0 or not (np.all(self.nu % n == 0) and np.all(self.nv % n == 0)):
            msg = "'n' should be a positive integer that divides all detector dimensions."
            raise ValueError(msg)
        return ProjGeom(source_pos=self.source_pos, det_pos=self.det_pos,
                        det_uvec=self.det_uvec * n, det_vvec=self.det_vvec * n,
                        nu=self.nu // n, nv=self.nv // n)

    def __eq__(self, other) -> bool:
        if not isinstance(other, ProjGeom):
            return False
        pairs = [(self.source_pos, other.source_pos), (self.det_pos, other.det_pos),
                 (self.det_uvec, other.det_uvec), (self.det_vvec, other.det_vvec),
                 (self.nu, other.nu), (self.nv, other.nv)]
        return all(a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a, b) for a, b in pairs)

    def __ne__(self, other) -> bool:
        return not self.__eq__(other)

    __hash__ = None  # mutable container of arrays

    # ---- JSON (same keys and nesting as the reference's files) ----------
    def json(self) -> dict:
        return {"source_pos": self.source_pos.tolist(), "det_pos": self.det_pos.tolist(),
                "det_uvec": self.det_uvec.tolist(), "det_vvec": self.det_vvec.tolist(),
                "nu": self.nu.tolist(), "nv": self.nv.tolist()}

    def json_str(self) -> str:
        return json.dumps(self.json(), indent=2)

    @staticmethod
    def _from_json_dict(d: dict) -> ProjGeom:
        return ProjGeom(source_pos=np.array(d["source_pos"], np.float32),
                        det_pos=np.array(d["det_pos"], np.float32),
                        det_uvec=np.array(d["det_uvec"], np.float32),
                        det_vvec=np.array(d["det_vvec"], np.float32),
                        nu=np.array(d["nu"], np.uint32), nv=np.array(d["nv"], np.uint32),
                        _skip_regularisation=True)

    @staticmethod
    def load(fpath) -> ProjGeom:
        fpath = pathlib.Path(fpath)
        if not fpath.exists():
            raise FileNotFoundError(fpath)
        with open(fpath, encoding="utf-8") as f:
            return ProjGeom._from_json_dict(json.load(f))

    def save(self, fpath, *, allow_overwrite: bool = False) -> None:
        fpath = pathlib.Path(fpath)
        if not allow_overwrite and fpath.exists():
            raise FileExistsError(fpath)
        with open(fpath, "w+", encoding="utf-8") as f:
            f.write(self.json_str())

    # ---- kernel arguments ---------------------------------------------
    def to_proj_spec_array(self) -> np.ndarray:
        """Array of ``proj_spec`` structs with the projection-list shape."""
        shp = self.projs_list_shape
        out = np.zeros(shp, dtype=_abi.proj_spec)
        for name, src in (("sourcepos", self.source_pos),):
            b = np.broadcast_to(src, shp + (3,))
            out[name]["x"], out[name]["y"], out[name]["z"] = b[..., 0], b[..., 1], b[..., 2]
        ds = out["dspec"]
        for name, src in (("pos", self.det_pos), ("u", self.det_uvec), ("v", self.det_vvec)):
            b = np.broadcast_to(src, shp + (3,))
            ds[name]["x"], ds[name]["y"], ds[name]["z"] = b[..., 0], b[..., 1], b[..., 2]
        ds["nu"] = np.broadcast_to(self.nu, shp)
        ds["nv"] = np.broadcast_to(self.nv, shp)
        ds["type"] = 0
        out["dspec"] = ds
        return out

    @classmethod
    def new_circular_scan(cls, *, num_source_points: int, R: float, L_minus_R: float,  # noqa: N803
                          pix_widhei_at_isocenter: tuple[float, float],
                          det_widhei: tuple[float, float] | None = None,
                          num_pix_widhei: tuple[int, int] | None = None,
                          theta_start: float = 0, theta_clockwise: bool = False,
                          arc_length_rads: float = 2 * np.pi,
                          isocenter=(0, 0, 0)) -> ProjGeom:
        """Circular scan: source R(cos t, sin t, 0), detector at distance L=R+L_minus_R."""
        if (det_widhei is None) == (num_pix_widhei is None):
            msg = "Exactly one of 'det_widhei' and 'num_pix_widhei' must be specified."
            raise ValueError(msg)
        L = R + L_minus_R  # noqa: N806
        px, py = pix_widhei_at_isocenter
        pw, ph = px * L / R, py * L / R
        if num_pix_widhei is None:
            num_pix_widhei = (math.ceil(det_widhei[0] / pw), math.ceil(det_widhei[1] / ph))
        t = np.linspace(theta_start, theta_start + arc_length_rads * (-1 if theta_clockwise else 1),
                        num=num_source_points, endpoint=False, dtype=np.float64)
        src = np.stack([R * np.cos(t), R * np.sin(t), np.zeros_like(t)], axis=1)
        det = src * np.array([1 - L / R, 1 - L / R, 1.0])
        iso = np.asarray(isocenter, dtype=np.float64)[None, :]
        u = pw * np.stack([-np.sin(t), np.cos(t), np.zeros_like(t)], axis=1)
        v = ph * np.array([[0.0, 0.0, -1.0]])
        return cls(source_pos=src + iso, det_pos=det + iso, det_uvec=u, det_vvec=v,
                   nu=num_pix_widhei[0], nv=num_pix_widhei[1])


class ProjectionSet:
    """A ProjGeom together with (optionally) its projection data on the GPU."""

    def __init__(self, pg: ProjGeom, *, data=None, device=None) -> None:
        torch = _torch()
        specs = pg.to_proj_spec_array()
        if data is not None and not isinstance(data, torch.Tensor):
            arr = np.asarray(data)
            data = torch.from_numpy(np.ascontiguousarray(arr)).to(device or "cuda")
        _check_geometry(specs, data)
        self._proj_geom = pg
        self._proj_specs = specs
        self._data = data
        self._device = data.device if data is not None else torch.device(device or "cuda")
        self._specs_dev = None
        self._out_of_mem_counter = None

    @property
    def proj_geom(self) -> ProjGeom:
        return self._proj_geom

    def get_bbox(self, *, index_tuple=None):
        return self._proj_geom.get_bbox(index_tuple=index_tuple)

    @property
    def proj_specs(self) -> np.ndarray:
        return self._proj_specs

    @property
    def data(self):
        """The underlying torch tensor ``[..., nu, nv]`` (or None)."""
        return self._data

    @property
    def device(self):
        return self._device

    def specs_device(self):
        """proj_spec array uploaded once to the GPU (uint8 tensor of raw structs)."""
        torch = _torch()
        if self._specs_dev is None:
            raw = np.ascontiguousarray(self._proj_specs.reshape(-1)).view(np.uint8)
            self._specs_dev = torch.from_numpy(raw.copy()).to(self._device)
        return self._specs_dev

    @property
    def out_of_mem_counter(self):
        torch = _torch()
        if self._out_of_mem_counter is None:
            self._out_of_mem_counter = torch.zeros(1, dtype=torch.int32, device=self._device)
        return self._out_of_mem_counter

    @property
    def out_of_mem_count(self) -> int:
        return int(self.out_of_mem_counter.item())

    def create_buffer(self, *, dtype=np.float32, existing_ok: bool = False) -> None:
        torch = _torch()
        if self._data is not None:
            if existing_ok:
                return
            msg = "Data buffer already exists. Pass 'existing_ok=True' to allow this to occur without raising an exception."
            raise RuntimeError(msg)
        shape = _check_geometry(self._proj_specs, None)
        self._data = torch.zeros(shape, dtype=_np_to_torch_dtype(dtype), device=self._device)

    def zeroed_copy(self, dtype=None) -> ProjectionSet:
        torch = _torch()
        if self._data is None and dtype is None:
            msg = "Cannot create a zeroed copy without an explicit 'dtype' argument, because the object you are trying to copy does not yet have a determined dtype."
            raise ValueError(msg)
        if self._data is None:
            shape = _check_geometry(self._proj_specs, None)
            data = torch.zeros(shape, dtype=_np_to_torch_dtype(dtype), device=self._device)
        else:
            data = torch.zeros_like(self._data)
        return ProjectionSet(self._proj_geom, data=data)

    # flattened (V, nu, nv) view for the kernels
    def _flat_data(self):
        d = self._data
        if d.dim() == 3:
            return d
        try:
            return d.view(-1, d.shape[-2], d.shape[-1])
        except RuntimeError as e:
            msg = "projection buffers with non-mergeable leading axes are not supported"
            raise NotImplementedError(msg) from e


class ConfigProject:
    """Projection settings (``config_CB_project`` / ``config_CB_project_FORBILD``)."""

    def __init__(self, *, shrink_volume_factor=None, multiply_before_adding=None,
                 ssu: int | None = None, ssv: int | None = None) -> None:
        # falsy values mean "default" exactly as in cb/flatpanel.py:903-906
        svf = shrink_volume_factor or 1
        mba = multiply_before_adding or 1
        ssu = ssu or 1
        ssv = ssv or 1
        self._contents = np.array((svf, mba, ssu, ssv), dtype=_abi.config_cb_project)
        self._contents_fb = np.array((mba, False, ssu, ssv), dtype=_abi.config_cb_project_FORBILD)

    @property
    def contents(self) -> np.ndarray:
        return self._contents.copy()

    @property
    def contents_fb(self) -> np.ndarray:
        return self._contents_fb.copy()


class ConfigBackproject:
    """Backprojection settings (``config_CB_backproject``)."""

    def __init__(self, *, fast_approx: bool = False, cyl_mask_rad=None, backproject_1s: bool = False,
                 apply_summand_multiplier: bool = False, soften_det_edges_proportion=None,
                 Z_on_L_factor_power: int = 0, Z_on_val_factor_power: int = 0,  # noqa: N803
                 Z_on_val_factor_val: float = 1, approx_mode_direct_raycast: bool = False,  # noqa: N803
                 approx_mode_bilinear_interpolate: bool = False, backproj_weighting_preset: int = 0,
                 ssx: int = 1, ssy: int = 1, ssz: int = 1, soft_vartheta_v=None,
                 soft_vartheta_v_softspan=None, soft_vartheta_h=None,
                 soft_vartheta_h_softspan=None) -> None:
        r = cyl_mask_rad or 0
        sep = soften_det_edges_proportion or 0
        self._contents = np.array(
            (not fast_approx, backproject_1s, bool(r), r, apply_summand_multiplier, sep,
             Z_on_L_factor_power, Z_on_val_factor_power, Z_on_val_factor_val,
             approx_mode_direct_raycast, approx_mode_bilinear_interpolate, backproj_weighting_preset,
             ssx, ssy, ssz, soft_vartheta_v or 0, soft_vartheta_v_softspan or 0,
             soft_vartheta_h or 0, soft_vartheta_h_softspan or 0),
            dtype=_abi.config_cb_backproject,
        )

    @property
    def contents(self) -> np.ndarray:
        return self._contents.copy()


# ==========================================================================
# operators
# ==========================================================================
def _chunk_of(volume: Volume) -> np.ndarray:
    texpos = volume.pos_c.astype(np.float64) - volume.voxwid * np.array(volume.shape, np.float64) / 2
    return np.array((volume.voxwid, tuple(0 for _ in range(3)), tuple(volume.shape), tuple(texpos)),
                    dtype=_abi.vol_tex_chunk_spec)


def project(config: ConfigProject, volume: Volume | FBVolume, projs: ProjectionSet,
            block: bool = False) -> ProjectionSet:
    """Add the projection of ``volume`` into ``projs.data`` (created as fp32 zeros if absent)."""
    torch = _torch()
    projs.create_buffer(dtype=np.float32, existing_ok=True)
    data = projs._flat_data()
    _abi.require_device(data, "projection data")
    specs = projs.specs_device()
    nprojs = data.shape[0]
    lib = _abi.lib()
    stream = _abi.stream_handle(data.device)
    sd = _abi.array3(data)
    if type(volume) is Volume:
        vdat = volume.data
        _abi.require_device(vdat, "volume")
        cfg = config.contents
        chunk = _chunk_of(volume)
        vd = _abi.array3(vdat)
        _abi.check(lib.ap_project_voxel(_abi.struct_ptr(cfg), _abi.ptr(specs), nprojs,
                                        _abi.struct_ptr(chunk), vd, _abi.ptr(vdat), sd, _abi.ptr(data),
                                        None, None, stream), "project")
    elif type(volume) is FBVolume:
        pruned = volume.prune_to_bbox(*projs.get_bbox())
        shapes = pruned.shapes_device(data.device)
        cfg = config.contents_fb
        counter = projs.out_of_mem_counter
        _abi.check(lib.ap_project_forbild(_abi.struct_ptr(cfg), _abi.ptr(specs), nprojs,
                                          _abi.ptr(shapes), len(pruned.shapes), sd, _abi.ptr(data),
                                          _abi.ptr(counter), stream), "project (FORBILD)")
    else:
        msg = f"Unknown volume type: {type(volume)}"
        raise RuntimeError(msg)
    if block:
        torch.cuda.synchronize(data.device)
    return projs


def backproject(config: ConfigBackproject, volume: Volume, projs: ProjectionSet, block: bool = False,
                summand_multiplier: Volume | None = None, logger=None, enable_pbar: bool = False,
                secondary_volume: Volume | None = None) -> Volume:
    """Add the backprojection of ``projs.data`` into ``volume`` (reference semantics)."""
    torch = _torch()
    del logger, enable_pbar  # progress reporting is not needed: one launch per call
    if projs.data is None:
        msg = "You must provide projection data to this method."
        raise ValueError(msg)
    if summand_multiplier is not None:
        if summand_multiplier.shape != volume.shape:
            msg = f"Shape mismatch between volume and multiplier. Volume shape: {volume.shape}. Multiplier shape: {summand_multiplier.shape}"
            raise ValueError(msg)
        if summand_multiplier.dtype != volume.dtype:
            msg = "dtype mismatch between summand multiplier and the volume."
            raise NotImplementedError(msg)
    cfg = config.contents
    if bool(cfg["apply_summand_multiplier"]) != (summand_multiplier is not None):
        msg = "To prevent unintended behaviour, the config 'apply_summand_multiplier' must be true if and only if a summand_multiplier is included."
        raise ValueError(msg)
    sdat = None
    if int(cfg["backproj_weighting_preset"]) == 7:
        # the GBC kernel also accumulates its weights (CB_backproject.cu:1033-1108)
        if secondary_volume is None:
            msg = "backproj_weighting_preset 7 (GBC) requires a secondary_volume for the sum of the weights."
            raise ValueError(msg)
        if secondary_volume.shape != volume.shape or secondary_volume.dtype != volume.dtype:
            msg = "The secondary volume must have the volume's shape and dtype."
            raise ValueError(msg)
        sdat = secondary_volume.data
        _abi.require_device(sdat, "secondary volume")
    data = projs._flat_data()
    vdat = volume.data
    _abi.require_device(data, "projection data")
    _abi.require_device(vdat, "volume")
    lib = _abi.lib()
    chunk = _chunk_of(volume)
    mdat = summand_multiplier.data if summand_multiplier is not None else None
    md = _abi.array3(mdat) if mdat is not None else None
    sd = _abi.array3(sdat) if sdat is not None else None
    _abi.check(lib.ap_backproject(_abi.struct_ptr(cfg), _abi.ptr(projs.specs_device()), data.shape[0],
                                  _abi.struct_ptr(chunk), _abi.array3(vdat), _abi.ptr(vdat), md,
                                  _abi.ptr(mdat), _abi.array3(data), _abi.ptr(data), sd, _abi.ptr(sdat),
                                  _abi.stream_handle(data.device)), "backproject")
    if block:
        torch.cuda.synchronize(data.device)
    return volume


def backproject_matched(config: ConfigProject, volume: Volume, projs: ProjectionSet, *,
                        alpha: float = 1.0, summand_multiplier: Volume | None = None,
                        block: bool = False) -> Volume:
    """``volume += alpha * (multiplier *) A^T projs`` with A exactly ``project(config, ., .)``.

    The transpose of the Siddon forward operator (same weights, same
    sub-pixel averaging and multiply_before_adding).  At ssu=ssv=1 it is the
    operator of ``backproject(ConfigBackproject(fast_approx=False), ...)``.
    """
    torch = _torch()
    if projs.data is None:
        msg = "You must provide projection data to this method."
        raise ValueError(msg)
    data = projs._flat_data()
    vdat = volume.data
    _abi.require_device(data, "projection data")
    _abi.require_device(vdat, "volume")
    mdat = summand_multiplier.data if summand_multiplier is not None else None
    md = _abi.array3(mdat) if mdat is not None else None
    cfg = config.contents
    chunk = _chunk_of(volume)
    ws, wsb = _abi.matched_workspace(cfg, data.shape[0], data)
    _abi.check(_abi.lib().ap_backproject_matched(
        _abi.struct_ptr(cfg), _abi.ptr(projs.specs_device()), data.shape[0], _abi.struct_ptr(chunk),
        _abi.array3(vdat), _abi.ptr(vdat), md, _abi.ptr(mdat), _abi.array3(data), _abi.ptr(data),
        float(alpha), ws, wsb, _abi.stream_handle(data.device)), "backproject_matched")
    if block:
        torch.cuda.synchronize(data.device)
    return volume


# typing helper for callers
ProjectionData = typing.Any
