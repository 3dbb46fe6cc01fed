"""Benchmark and test workloads: the geometries of the target configurations.

Choices the configurations leave open are fixed here and recorded:
  * circular scans use loadProjectionGeometryCT's angles ((1:N) - N/2) * 2*pi/N;
  * PWLS-OS 512^2 x 128: SAD 1000 mm, SDD 1500 mm, flat detector 1024 x 64 at
    0.75 mm x 3.0 mm (covers the 358 mm volume, 192 mm rows at isocentre), FOV map of
    loadProjectionGeometryCT.m:42-53;
  * attenuation in /mm is multiplied by 10 into the reference's 1/cm convention.
"""
from __future__ import annotations

import math

import numpy as np

from .geometry import Geometry, circular_betas, fov_map


def sf_pair_512():
    """3D cone-beam SF pair: 512x512x256 at 0.5 mm, 720 views, 1024x384 at 0.4 mm."""
    return Geometry(reconSize=(512, 512, 256), reconSpacing=(0.5, 0.5, 0.5), reconOffset=(0, 0, 0),
                    detSize=(1024, 384), detSpacing=(0.4, 0.4), detOffset=(0, 0), SAD=1000.0, ADD=500.0,
                    betas=circular_betas(720), flatPanel=1)


def pwls_512x128(with_map=True):
    """PWLS-OS volume 512x512x128 (0.7 mm, 1.0 mm slices), 984 views, 1024x64 detector."""
    g = Geometry(reconSize=(512, 512, 128), reconSpacing=(0.7, 0.7, 1.0), reconOffset=(0, 0, 0),
                 detSize=(1024, 64), detSpacing=(0.75, 3.0), detOffset=(0, 0), SAD=1000.0, ADD=500.0,
                 betas=circular_betas(984), flatPanel=1)
    if with_map:
        g.map = fov_map(g)
    return g


def fdk_512():
    """FDK: 720 views, flat 1024x768 at 0.388 mm, 512^3 volume at 0.5 mm."""
    return Geometry(reconSize=(512, 512, 512), reconSpacing=(0.5, 0.5, 0.5), reconOffset=(0, 0, 0),
                    detSize=(1024, 768), detSpacing=(0.388, 0.388), detOffset=(0, 0), SAD=1000.0, ADD=500.0,
                    betas=circular_betas(720), flatPanel=1)


def fan_256():
    """2D DD pair: 256^2 at 0.8 mm, 360 views over 360 deg, 512 bins at 1.0 mm."""
    return Geometry(reconSize=(256, 256), reconSpacing=(0.8, 0.8), reconOffset=(0, 0), detSize=(512,),
                    detSpacing=(1.0,), detOffset=(0,), SAD=1000.0, ADD=500.0, betas=circular_betas(360),
                    flatPanel=1)


def ellipsoid_phantom(shape, spacing, n, seed, semi=(5.0, 120.0), mu=(0.015, 0.06), dtype=np.float32):
    """n random ellipsoids (additive), centres inside the inscribed cylinder, random
    rotation about z; mu in /mm times 10 (1/cm).  Returns MATLAB-order [nx ny nz]."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    dx, dy, dz = spacing
    x = (np.arange(nx) - (nx - 1) / 2) * dx
    y = (np.arange(ny) - (ny - 1) / 2) * dy
    z = (np.arange(nz) - (nz - 1) / 2) * dz
    R = 0.45 * min(nx * dx, ny * dy)
    H = 0.45 * nz * dz
    vol = np.zeros(shape, np.float32)
    X, Y = np.meshgrid(x, y, indexing="ij")
    for _ in range(n):
        r = R * np.sqrt(rng.uniform()) * 0.8
        ph = rng.uniform(0, 2 * np.pi)
        cx, cy, cz = r * np.cos(ph), r * np.sin(ph), rng.uniform(-0.6, 0.6) * H
        a, b, c = rng.uniform(*semi, 3)
        a, b = min(a, R), min(b, R)
        c = min(c, H)
        th = rng.uniform(0, np.pi)
        val = np.float32(rng.uniform(*mu) * 10.0)
        ct, st = np.cos(th), np.sin(th)
        xr = (X - cx) * ct + (Y - cy) * st
        yr = -(X - cx) * st + (Y - cy) * ct
        inplane = (xr / a) ** 2 + (yr / b) ** 2
        for k in range(nz):
            if abs(z[k] - cz) > c:
                continue
            q = inplane + ((z[k] - cz) / c) ** 2
            m = q <= 1.0
            if m.any():
                vol[:, :, k][m] += val
    return vol.astype(dtype)


def body_phantom(shape, spacing, n_inserts, seed, water=0.02, mu=(0.015, 0.06), dtype=np.float32):
    """Water cylinder (radius 40% of the field, mu 0.02 /mm) with n random ellipsoidal
    inserts (semi-axes 5-60 mm, mu 0.015-0.06 /mm) that replace the water, as
    phantom/util/addEllipses.m does; values x10 into 1/cm.  Line integrals stay below ~8,
    so rays keep tens of photons at I0 = 2e4.  MATLAB order [nx ny nz]."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    dx, dy, dz = spacing
    x = (np.arange(nx) - (nx - 1) / 2) * dx
    y = (np.arange(ny) - (ny - 1) / 2) * dy
    z = (np.arange(nz) - (nz - 1) / 2) * dz
    X, Y = np.meshgrid(x, y, indexing="ij")
    R = 0.4 * min(nx * dx, ny * dy)
    H = 0.5 * nz * dz
    slab = ((X / R) ** 2 + (Y / R) ** 2 <= 1.0).astype(np.float32) * np.float32(water * 10.0)
    vol = np.repeat(slab[:, :, None], nz, axis=2)
    for _ in range(n_inserts):
        r = R * 0.7 * np.sqrt(rng.uniform())
        ph = rng.uniform(0, 2 * np.pi)
        cx, cy, cz = r * np.cos(ph), r * np.sin(ph), rng.uniform(-0.5, 0.5) * H
        a, b = rng.uniform(5.0, 60.0, 2)
        c = rng.uniform(5.0, 60.0)
        th = rng.uniform(0, np.pi)
        val = np.float32(rng.uniform(*mu) * 10.0)
        ct, st = np.cos(th), np.sin(th)
        xr = (X - cx) * ct + (Y - cy) * st
        yr = -(X - cx) * st + (Y - cy) * ct
        inplane = (xr / a) ** 2 + (yr / b) ** 2
        for k in range(nz):
            if abs(z[k] - cz) > c:
                continue
            m = inplane + ((z[k] - cz) / c) ** 2 <= 1.0
            vol[:, :, k][m] = val
    return vol.astype(dtype)


def batch2_uneven(n_slices=64, seed=5):
    """64 independent fan-beam slices: image sizes from {128, 256, 384, 512}^2 at 0.8 mm,
    view counts from {180, 360, 720, 984} over 360 degrees, SAD 1000 / SDD 1500 mm, flat
    detector at 1.0 mm with 256-1024 bins covering the field of view (the image diagonal
    magnified by SDD/SAD, rounded up to a multiple of 64).  Each slice holds 10 random
    ellipses from its own seed (ellipse_image).  Returns (geometries, images)."""
    rng = np.random.default_rng(seed)
    geoms, imgs = [], []
    for i in range(n_slices):
        n = int(rng.choice([128, 256, 384, 512]))
        nviews = int(rng.choice([180, 360, 720, 984]))
        width = n * 0.8 * math.sqrt(2.0) * 1.5
        nu = int(min(1024, max(256, 64 * math.ceil(width / 64.0))))
        g = Geometry(reconSize=(n, n), reconSpacing=(0.8, 0.8), reconOffset=(0, 0), detSize=(nu,),
                     detSpacing=(1.0,), detOffset=(0,), SAD=1000.0, ADD=500.0, betas=circular_betas(nviews),
                     flatPanel=1)
        geoms.append(g)
        imgs.append(ellipse_image(n, 0.8, 10, seed * 1000 + i))
    return geoms, imgs


def ellipse_image(n, dx, n_ell, seed, semi=(5.0, 90.0), mu=(0.01, 0.05)):
    """n x n image of n_ell random ellipses (centres inside the field, semi-axes in mm,
    random rotation, attenuation in 1/mm converted to the reference's 1/cm), additive
    values as in phantom/util/ellipses.m:35-36."""
    rng = np.random.default_rng(seed)
    x = (np.arange(n) - (n - 1) / 2) * dx
    X, Y = np.meshgrid(x, x, indexing="ij")
    img = np.zeros((n, n), np.float64)
    rmax = 0.45 * n * dx
    for _ in range(n_ell):
        r = rmax * math.sqrt(rng.uniform())
        t = rng.uniform(0, 2 * math.pi)
        cx, cy = r * math.cos(t), r * math.sin(t)
        a, b = rng.uniform(*semi, size=2)
        ang = rng.uniform(0, math.pi)
        ca, sa = math.cos(ang), math.sin(ang)
        xr = (X - cx) * ca + (Y - cy) * sa
        yr = -(X - cx) * sa + (Y - cy) * ca
        img[(xr / a) ** 2 + (yr / b) ** 2 <= 1.0] += rng.uniform(*mu) * 10.0
    return img.astype(np.float32)
