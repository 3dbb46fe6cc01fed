"""Shared builders for tests (geometry from golden records, phantoms)."""
import numpy as np

from oracle import oracle as O


def g2_from(rec):
    return O.G2(int(rec["nxy"][0]), int(rec["nxy"][1]), float(rec["dxy"][0]), float(rec["dxy"][1]),
                float(rec["oxy"][0]), float(rec["oxy"][1]), int(rec["nu"]), float(rec["du"]), float(rec["ou"]),
                float(rec["sad"]), float(rec["add"]), float(rec["sdd"]), rec["betas"], int(rec["dettype"]))


def g3_from(rec):
    return O.G3(rec["n"], rec["d"], rec["o"], rec["nuv"], rec["duv"], rec["ouv"], float(rec["sad"]),
                float(rec["add"]), float(rec["sdd"]), rec["betas"], couchz=rec["couchz"], dettype=int(rec["dettype"]))


def geometry_from_g2(g):
    from ctsim_amd.geometry import Geometry
    return Geometry(reconSize=g.nxy, reconSpacing=g.dxy, reconOffset=g.oxy, detSize=(g.nu,), detSpacing=(g.du,),
                    detOffset=(g.ou,), SAD=g.sad, ADD=g.sdd - g.sad, SDD=g.sdd, betas=g.betas, flatPanel=g.dettype)


def geometry_from_g3(g, map_=None):
    from ctsim_amd.geometry import Geometry
    return Geometry(reconSize=g.n, reconSpacing=g.d, reconOffset=g.o, detSize=g.nuv, detSpacing=g.duv,
                    detOffset=g.ouv, SAD=g.sad, ADD=g.sdd - g.sad, SDD=g.sdd, betas=g.betas, flatPanel=g.dettype,
                    couchZ=g.couchz, map=map_)


def ellipse_phantom2(nx, ny, dx, n_ell, rng, semi=(5.0, 90.0), mu=(0.01, 0.05)):
    """Random ellipses (additive, as phantom/util/ellipses.m:35-36), mu in /mm converted to the
    reference's 1/cm convention (x10)."""
    x = (np.arange(nx) - (nx - 1) / 2) * dx
    y = (np.arange(ny) - (ny - 1) / 2) * dx
    X, Y = np.meshgrid(x, y, indexing="ij")
    img = np.zeros((nx, ny), np.float64)
    rmax = 0.45 * min(nx, ny) * dx
    for _ in range(n_ell):
        cx, cy = rng.uniform(-0.5, 0.5, 2) * rmax
        a, b = rng.uniform(*semi, 2) * (rmax / 128.0 if rmax < 128 else 1.0)
        th = rng.uniform(0, np.pi)
        c, s = np.cos(th), np.sin(th)
        xr = (X - cx) * c + (Y - cy) * s
        yr = -(X - cx) * s + (Y - cy) * c
        img[(xr / a) ** 2 + (yr / b) ** 2 <= 1] += rng.uniform(*mu) * 10.0
    return img.astype(np.float32)
