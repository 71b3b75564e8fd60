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
}


def voxel_config(i: int):
    """(geometry, shapes, N, voxwid) of voxel workload ``i``."""
    c = CONFIGS[i]
    pg = circular_geometry(c["views"], c["nu"], c["nv"], c["pitch"])
    shapes = random_ellipsoids(c["n_ellipsoids"], c["N"] * c["voxwid"] / 2, c["seed"])
    return pg, shapes, c["N"], c["voxwid"]
