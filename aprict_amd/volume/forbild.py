"""Constructors for FORBILD primitive structs (``FORBILD_Shape``).

Type ids and field meanings follow ``src/aprict/cuda/structs/FORBILD_volumes.py``
(:502-607): Box ``dx,dy,dz`` are full side lengths; ellipsoid semi-axes
``dx,dy,dz`` lie along the normalised ``a_x,a_y,a_z``; elliptic cylinders have
semi-axes ``dx,dy`` along ``a_x,a_y`` and full length ``l`` along ``axis``;
cones have radius ``r1`` at -l/2 and ``r2`` at +l/2 along their axis
(``cuda/structs/cu/FORBILD_Cone_z.cu:1-13``); clip plane mode 1 keeps
``axis.r <= d``, mode 2 keeps ``axis.r > d``.
"""

from __future__ import annotations

import numpy as np

from .. import _abi

TYPE_IDS = {
    "Box": 1, "Cone": 2, "Cone_x": 3, "Cone_y": 4, "Cone_z": 5, "Cylinder": 6, "Cylinder_x": 7,
    "Cylinder_y": 8, "Cylinder_z": 9, "Ellipsoid": 10, "Ellipsoid_free": 11, "Ellipt_Cyl": 12,
    "Sphere": 13, "Ellipt_Cyl_x": 14, "Ellipt_Cyl_y": 15, "Ellipt_Cyl_z": 16,
}
SUPPORTED = set(range(1, 17))


def _v(a) -> tuple[float, float, float]:
    a = np.asarray(a, dtype=np.float64).reshape(3)
    return float(a[0]), float(a[1]), float(a[2])


def empty_shape() -> np.ndarray:
    return np.zeros((), dtype=_abi.forbild_shape)


def _set_clips(s: np.ndarray, clips) -> None:
    clips = list(clips or [])
    if len(clips) > 3:
        msg = "at most 3 clipping planes per shape"
        raise ValueError(msg)
    for n, (axis, d, mode) in enumerate(clips):
        c = s[f"cas{n + 1}"]
        c["axis"] = _v(axis)
        c["dotprod"] = d
        c["mode"] = mode
        s[f"cas{n + 1}"] = c


def shape(type_name: str, *, rho: float, center=(0, 0, 0), clips=None, **fields) -> np.ndarray:
    """One FORBILD_Shape struct; ``fields`` are dx, dy, dz, r, l, a_x, a_y, a_z, axis."""
    s = empty_shape()
    s["type"] = TYPE_IDS[type_name]
    s["rho"] = rho
    s["x"], s["y"], s["z"] = _v(center)
    for k, val in fields.items():
        if k in ("a_x", "a_y", "a_z", "axis"):
            s[k] = _v(val)
        else:
            s[k] = val
    _set_clips(s, clips)
    return s


def box(center, dx, dy, dz, rho, clips=None):
    return shape("Box", rho=rho, center=center, dx=dx, dy=dy, dz=dz, clips=clips)


def sphere(center, r, rho, clips=None):
    return shape("Sphere", rho=rho, center=center, r=r, clips=clips)


def ellipsoid(center, dx, dy, dz, rho, clips=None):
    return shape("Ellipsoid", rho=rho, center=center, dx=dx, dy=dy, dz=dz, clips=clips)


def ellipsoid_free(center, semi_axes, axes, rho, clips=None):
    """Ellipsoid with semi-axes ``semi_axes`` along the rows ``a_x, a_y, a_z`` of ``axes``."""
    ax = np.asarray(axes, dtype=np.float64).reshape(3, 3)
    return shape("Ellipsoid_free", rho=rho, center=center, dx=semi_axes[0], dy=semi_axes[1],
                 dz=semi_axes[2], a_x=ax[0], a_y=ax[1], a_z=ax[2], clips=clips)


def ellipt_cyl(center, dx, dy, l, a_x, a_y, axis, rho, clips=None):  # noqa: E741
    return shape("Ellipt_Cyl", rho=rho, center=center, dx=dx, dy=dy, l=l, a_x=a_x, a_y=a_y,
                 axis=axis, clips=clips)


def cylinder(center, r, l, axis, rho, clips=None):  # noqa: E741
    return shape("Cylinder", rho=rho, center=center, r=r, l=l, axis=axis, clips=clips)


def cone(center, r1, r2, l, axis, rho, clips=None):  # noqa: E741
    """Truncated cone of length ``l``: radius ``r1`` at -l/2 and ``r2`` at +l/2
    along ``axis`` ("x", "y", "z" or a vector)."""
    if isinstance(axis, str):
        return shape({"x": "Cone_x", "y": "Cone_y", "z": "Cone_z"}[axis], rho=rho, center=center, r1=r1, r2=r2,
                     l=l, clips=clips)
    return shape("Cone", rho=rho, center=center, r1=r1, r2=r2, l=l, axis=axis, clips=clips)


def stack(shapes) -> np.ndarray:
    return np.array([np.asarray(s) for s in shapes], dtype=_abi.forbild_shape)
