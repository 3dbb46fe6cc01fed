"""Writes the golden fixtures of tests/golden/ from the reference's own CPU kernels.

The reference kernels (geom/fbct_os/*.c, geom/cbct_os/*.c) are compiled from their
sources under /root/reference by oracle/Makefile into oracle/_ref/libctsim_ref.so; this
script calls them on small seeded cases and stores inputs and outputs as .npz.  The
fixtures are data only (inputs, expected outputs, geometry parameters).

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def cases2():
    rng = np.random.default_rng(20240601)
    out = []
    # (name, nx, ny, dx, dy, ox, oy, nu, du, ou, sad, sdd, nviews, dettype)
    specs = [
        ("fan_flat_small", 24, 24, 0.8, 0.8, 0.0, 0.0, 48, 1.0, 0.0, 1000.0, 1500.0, 16, 1),
        ("fan_flat_offsets", 20, 28, 0.7, 0.9, 1.5, -2.0, 40, 1.1, 2.25, 600.0, 1000.0, 12, 1),
        ("fan_arc", 22, 22, 1.0, 1.0, 0.0, 0.5, 44, 1.0, -1.5, 541.0, 949.0, 10, 0),
        ("fan_fine_detector", 16, 16, 2.0, 2.0, 0.0, 0.0, 64, 0.25, 0.0, 500.0, 800.0, 9, 1),
    ]
    for name, nx, ny, dx, dy, ox, oy, nu, du, ou, sad, sdd, nv, dt in specs:
        betas = np.sort(rng.uniform(0, 2 * np.pi, nv))
        g = O.G2(nx, ny, dx, dy, ox, oy, nu, du, ou, sad, 0.0, sdd, betas, dt)
        img = rng.uniform(0, 1, (nx, ny)).astype(np.float32)
        img[rng.uniform(size=img.shape) < 0.15] = 0
        sino = rng.uniform(0, 1, (nu, nv)).astype(np.float32)
        rec = dict(nxy=g.nxy, dxy=g.dxy, oxy=g.oxy, nu=np.int32(nu), du=np.float32(du), ou=np.float32(ou),
                   sad=np.float32(sad), add=np.float32(0), sdd=np.float32(sdd), betas=betas, dettype=np.int32(dt),
                   img=img, sino=sino)
        for op in ("proj,rd", "proj,dd", "proj,tf"):
            rec[op.replace(",", "_")] = O.ref_fbct(op, g, img)
        for op in ("back,dd", "back,tf", "back,pd"):
            rec[op.replace(",", "_")] = O.ref_fbct(op, g, sino)
        out.append((name, rec))
    return out


def cases3():
    rng = np.random.default_rng(20240602)
    out = []
    # (name, n, d, o, nuv, duv, ouv, sad, sdd, nviews, dettype, use_map, couch)
    specs = [
        ("cone_flat_map", (14, 12, 10), (0.5, 0.5, 0.5), (0, 0, 0), (40, 24), (0.4, 0.4), (0, 0), 1000.0, 1500.0,
         8, 1, True, False),
        ("cone_flat_nomap_offsets", (12, 14, 9), (0.7, 0.6, 0.8), (0.5, -1.0, 0.75), (36, 20), (0.5, 0.6),
         (1.5, -2.0), 700.0, 1100.0, 6, 1, False, False),
        ("cone_couch", (10, 10, 12), (1.0, 1.0, 1.0), (0, 0, 0), (30, 28), (0.8, 0.7), (0, 0.5), 500.0, 900.0, 6,
         1, True, True),
        ("cone_arc", (12, 12, 8), (0.9, 0.9, 1.1), (0, 0, 0), (36, 16), (0.6, 0.9), (-0.5, 0), 541.0, 949.0, 5,
         0, True, False),
        ("cone_coarse_detector", (16, 16, 6), (0.4, 0.4, 1.5), (0, 0, 0), (20, 8), (1.2, 2.5), (0, 0), 800.0,
         1200.0, 7, 1, False, False),
    ]
    for name, n, d, o, nuv, duv, ouv, sad, sdd, nv, dt, use_map, couch in specs:
        betas = np.sort(rng.uniform(0, 2 * np.pi, nv))
        cz = rng.uniform(-2, 2, nv).astype(np.float32) if couch else np.zeros(nv, np.float32)
        g = O.G3(n, d, o, nuv, duv, ouv, sad, 0.0, sdd, betas, couchz=cz, dettype=dt)
        vol = rng.uniform(0, 1, n).astype(np.float32)
        vol[rng.uniform(size=vol.shape) < 0.15] = 0
        sino = rng.uniform(0, 1, (nuv[1], nuv[0], nv)).astype(np.float32)
        mp = (rng.uniform(size=n[:2]) < 0.85) if use_map else None
        rec = dict(n=g.n, d=g.d, o=g.o, nuv=g.nuv, duv=g.duv, ouv=g.ouv, sad=np.float32(sad), add=np.float32(0),
                   sdd=np.float32(sdd), betas=betas, couchz=cz, dettype=np.int32(dt), img=vol, sino=sino,
                   has_map=np.bool_(use_map), map=(mp if mp is not None else np.ones(n[:2], bool)))
        for op in ("proj,dd", "proj,tf"):
            rec[op.replace(",", "_")] = O.ref_cbct(op, g, vol, mp)
        for op in ("back,dd", "back,tf", "back,pd", "back,hpd"):
            rec[op.replace(",", "_")] = O.ref_cbct(op, g, sino, mp)
        # the ray-driven projector exists only in the 15-argument (no map) family
        rec["proj_rd"] = O.ref_cbct("proj,rd", g, vol, None)
        out.append((name, rec))
    return out


def main():
    O.build(ref=True)
    if not O.ref_available():
        raise SystemExit("oracle/_ref is not built (reference tree missing)")
    for name, rec in cases2() + cases3():
        path = os.path.join(OUT, f"{name}.npz")
        np.savez_compressed(path, **rec)
        print("wrote", os.path.relpath(path, ROOT))


if __name__ == "__main__":
    main()
