"""Small problem builders shared by the CPU and GPU tests."""

from __future__ import annotations

import numpy as np

from aprict_amd import _abi
from aprict_amd.cb import flatpanel as cbfp


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def chunk_for(shape, voxwid, pos_c=(0.0, 0.0, 0.0), offsets=(0, 0, 0), cshape=None) -> np.ndarray:
    shape = tuple(int(s) for s in shape)
    texpos = np.asarray(pos_c, np.float64) - voxwid * np.array(shape, np.float64) / 2
    return np.array((voxwid, tuple(offsets), tuple(cshape or shape), tuple(texpos)),
                    dtype=_abi.vol_tex_chunk_spec)


def random_geometry(rng, n_views: int, nu: int, nv: int, *, R=60.0, L=120.0, pitch=1.0,
                    tilt_deg=5.0, offset=3.0) -> cbfp.ProjGeom:
    """Arbitrary flat panels: random source directions (incl. out of plane),
    detector tilted about u and v, offset from the central ray."""
    src, det, us, vs = [], [], [], []
    for _ in range(n_views):
        d = rng.standard_normal(3)
        d[2] *= 0.4
        d /= np.linalg.norm(d)
        s = R * d
        c = s - L * d + rng.uniform(-offset, offset, 3)
        a = np.cross(d, [0.0, 0.0, 1.0])
        a /= np.linalg.norm(a)
        b = np.cross(d, a)
        t1, t2 = np.deg2rad(rng.uniform(-tilt_deg, tilt_deg, 2))
        u = np.cos(t1) * a + np.sin(t1) * d
        v = np.cross(d, u)
        v = np.cos(t2) * v + np.sin(t2) * np.cross(u, v)
        src.append(s)
        det.append(c)
        us.append(pitch * u)
        vs.append(pitch * v / np.linalg.norm(v))
        del b
    return cbfp.ProjGeom(source_pos=np.array(src), det_pos=np.array(det), det_uvec=np.array(us),
                         det_vvec=np.array(vs), nu=nu, nv=nv)


def pixel_point(spec, ju: float, kv: float) -> np.ndarray:
    d = spec["dspec"]
    p = np.array([d["pos"]["x"], d["pos"]["y"], d["pos"]["z"]], np.float64)
    u = np.array([d["u"]["x"], d["u"]["y"], d["u"]["z"]], np.float64)
    v = np.array([d["v"]["x"], d["v"]["y"], d["v"]["z"]], np.float64)
    return p + (ju - (int(d["nu"]) - 1) / 2) * u + (kv - (int(d["nv"]) - 1) / 2) * v


def source_point(spec) -> np.ndarray:
    s = spec["sourcepos"]
    return np.array([s["x"], s["y"], s["z"]], np.float64)
