"""Seeded synthetic inputs: circular scans, random ellipsoid and primitive phantoms.

Generators use ``numpy.random.default_rng(seed)`` (PCG64).  Random ellipsoids
are ``Ellipsoid_free`` shapes in last-wins order with, for a cube of
half-width H: centres U[-0.8,0.8]^3 H, semi-axes U[0.05,0.4]^3 H, a Haar
rotation (QR of a Gaussian matrix, signs fixed, det +1) whose rows are
a_x, a_y, a_z, and densities U[0,1].
"""

from __future__ import annotations

import numpy as np

from .cb import flatpanel as cbfp
from .forbild_presets import PhantomContainer as FORBILD  # noqa: F401  (aprict.phantoms.FORBILD)
from .volume import forbild


def circular_geometry(num_views: int, nu: int, nv: int, pitch: float, *, R: float = 500.0,  # noqa: N803
                      L: float = 1000.0) -> cbfp.ProjGeom:  # noqa: N803
    """Circular cone-beam scan, detector pitch ``pitch`` (mm) at distance L."""
    return cbfp.ProjGeom.new_circular_scan(num_source_points=num_views, R=R, L_minus_R=L - R,
                                           pix_widhei_at_isocenter=(pitch * R / L, pitch * R / L),
                                           num_pix_widhei=(nu, nv))


def haar_rotation(rng: np.random.Generator) -> np.ndarray:
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))[None, :]
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def random_ellipsoids(n: int, half_width: float, seed: int) -> np.ndarray:
    """``n`` Ellipsoid_free shapes inside a cube of half-width ``half_width``."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        c = rng.uniform(-0.8, 0.8, 3) * half_width
        ax = rng.uniform(0.05, 0.4, 3) * half_width
        rot = haar_rotation(rng)
        rho = rng.uniform(0.0, 1.0)
        out.append(forbild.ellipsoid_free(c, ax, rot.T, rho))
    return forbild.stack(out)


def random_primitives(n: int, scale: float, seed: int) -> np.ndarray:
    """FORBILD ellipsoids, elliptic cylinders and boxes with 0-3 clip planes.

    Types uniform over {Ellipsoid_free, Ellipt_Cyl, Box}; centres
    U[-0.8,0.8] scale; sizes U[0.05,0.4] scale (semi-axes; box sides and
    cylinder lengths are twice that); clip normals uniform on S^2 with offset
    n.c + U[-0.5,0.5] size; densities U[-0.5,1].
    """
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        kind = int(rng.integers(0, 3))
        c = rng.uniform(-0.8, 0.8, 3) * scale
        sz = rng.uniform(0.05, 0.4, 3) * scale
        rho = rng.uniform(-0.5, 1.0)
        nclip = int(rng.integers(0, 4))
        clips = []
        for _ in range(nclip):
            nrm = rng.standard_normal(3)
            nrm /= np.linalg.norm(nrm)
            d = float(nrm @ c + rng.uniform(-0.5, 0.5) * sz.mean())
            clips.append((nrm, d, int(rng.integers(1, 3))))
        if kind == 0:
            rot = haar_rotation(rng)
            out.append(forbild.ellipsoid_free(c, sz, rot.T, rho, clips=clips))
        elif kind == 1:
            rot = haar_rotation(rng)
            out.append(forbild.ellipt_cyl(c, sz[0], sz[1], 2 * sz[2], rot[:, 0], rot[:, 1], rot[:, 2], rho,
                                          clips=clips))
        else:
            out.append(forbild.box(c, 2 * sz[0], 2 * sz[1], 2 * sz[2], rho, clips=clips))
    return forbild.stack(out)


# The five workloads the path is measured and tested on.
CONFIGS = {
    1: {"N": 64, "voxwid": 0.5, "n_ellipsoids": 10, "views": 90, "nu": 96, "nv": 96, "pitch": 1.0, "seed": 1},
    2: {"N": 256, "voxwid": 0.5, "n_ellipsoids": 30, "views": 360, "nu": 384, "nv": 384, "pitch": 1.0,
        "seed": 2, "sirt_iters": 20},
    3: {"n_primitives": 40, "scale": 30.0, "views": 720, "nu": 512, "nv": 512, "pitch": 0.5, "seed": 3},
    4: {"N": 512, "voxwid": 0.48, "n_ellipsoids": 30, "views": 720, "nu": 768, "nv": 768, "pitch": 1.0,
        "seed": 4},
    5: {"problems": 64, "seed": 5, "sirt_iters": 10, "storage": "float16"},
}


def _rot(axis: np.ndarray, ang: float) -> np.ndarray:
    """Rotation matrix about the unit vector ``axis`` by ``ang`` radians (Rodrigues)."""
    a = axis / np.linalg.norm(axis)
    k = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(ang) * k + (1 - np.cos(ang)) * (k @ k)


def tilted_geometry(num_views: int, n: int, pitch: float, *, inplane_deg: float, outplane_deg: float,
                    offset_uv: tuple[float, float], R: float = 500.0, L: float = 1000.0) -> cbfp.ProjGeom:  # noqa: N803
    """Circular scan whose square ``n x n`` panel is tilted in its plane (about
    the detector normal) and out of it (about the u axis), and shifted by
    ``offset_uv`` mm along u and v; u and v stay orthogonal."""
    base = circular_geometry(num_views, n, n, pitch, R=R, L=L)
    src = base.source_pos.astype(np.float64)
    det = base.det_pos.astype(np.float64)
    u0 = base.det_uvec.astype(np.float64)
    v0 = np.broadcast_to(base.det_vvec.astype(np.float64), u0.shape)
    us, vs, ds = [], [], []
    for i in range(num_views):
        nrm = np.cross(u0[i], v0[i])
        r = _rot(u0[i], np.deg2rad(outplane_deg)) @ _rot(nrm, np.deg2rad(inplane_deg))
        u, v = r @ u0[i], r @ v0[i]
        us.append(u)
        vs.append(v)
        ds.append(det[i] + offset_uv[0] * u / np.linalg.norm(u) + offset_uv[1] * v / np.linalg.norm(v))
    return cbfp.ProjGeom(source_pos=src, det_pos=np.array(ds), det_uvec=np.array(us), det_vvec=np.array(vs),
                         nu=n, nv=n)


def batched_problems(count: int = 64, seed: int = 5, *, n_range=(32, 192), views_range=(60, 360),
                     det_range=(64, 256), pitch: float = 1.0, tilt_deg: float = 5.0, offset_mm: float = 10.0,
                     n_ellipsoids: int = 10, R: float = 500.0, L: float = 1000.0) -> list[dict]:  # noqa: N803
    """Uneven reconstruction problems: volume side N, view count V and square
    detector side n drawn uniformly; detector tilts U[-tilt, tilt] degrees in
    and out of its plane and centre offsets U[-offset, offset] mm along u and v;
    the voxel width is fitted so that the volume's circumscribed cylinder and
    its height project inside the detector.  Each entry holds the geometry,
    the FORBILD ellipsoid phantom and the sizes."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n_vol = int(rng.integers(n_range[0], n_range[1] + 1))
        views = int(rng.integers(views_range[0], views_range[1] + 1))
        n_det = int(rng.integers(det_range[0], det_range[1] + 1))
        tilt_in, tilt_out = rng.uniform(-tilt_deg, tilt_deg, 2)
        off = rng.uniform(-offset_mm, offset_mm, 2)
        # usable half-width at the detector, back-projected to the isocentre,
        # with room for the tilt and for magnification of the near side
        half_det = n_det * pitch / 2 - np.abs(off).max() - n_det * pitch * np.sin(np.deg2rad(tilt_deg))
        half_iso = 0.85 * half_det * R / L
        # circumscribed cylinder radius r = N w / sqrt(2); near-side magnification L / (R - r)
        w = half_iso * np.sqrt(2) / n_vol
        w = w / (1 + n_vol * w / np.sqrt(2) / R)
        pg = tilted_geometry(views, n_det, pitch, inplane_deg=float(tilt_in), outplane_deg=float(tilt_out),
                             offset_uv=(float(off[0]), float(off[1])), R=R, L=L)
        shapes = random_ellipsoids(n_ellipsoids, n_vol * w / 2, seed * 1000 + i)
        out.append({"geometry": pg, "shapes": shapes, "N": n_vol, "voxwid": float(w), "views": views,
                    "det": n_det, "tilt_deg": (float(tilt_in), float(tilt_out)), "offset_mm": tuple(off.tolist())})
    return out


def batched_problem_data(entry: dict, storage: str = "float16", device=None):
    """GPU data of one ``batched_problems`` entry: a zero initial volume and
    the measurements ``b = A x_true`` (x_true: the entry's ellipsoids point-
    sampled at voxel centres; A computed in FP32, then stored as ``storage``)."""
    import torch  # noqa: PLC0415

    from .volume import FBVolume, Volume  # noqa: PLC0415

    n, w = entry["N"], entry["voxwid"]
    truth = Volume.zeros((n, n, n), voxwid=w, dtype=np.float32, device=device)
    FBVolume(entry["shapes"]).voxelise(truth)
    ps = cbfp.ProjectionSet(entry["geometry"], device=device)
    cbfp.project(cbfp.ConfigProject(), truth, ps)
    tdt = {"float16": torch.float16, "float32": torch.float32}[storage]
    meas = cbfp.ProjectionSet(entry["geometry"], data=ps.data.to(tdt), device=device)
    x0 = Volume.zeros((n, n, n), voxwid=w, dtype=storage, device=device)
    return x0, meas, truth


def voxel_config(i: int):
    """(geometry, shapes, N, voxwid) of voxel workload ``i``."""
    c = CONFIGS[i]
    pg = circular_geometry(c["views"], c["nu"], c["nv"], c["pitch"])
    shapes = random_ellipsoids(c["n_ellipsoids"], c["N"] * c["voxwid"] / 2, c["seed"])
    return pg, shapes, c["N"], c["voxwid"]
