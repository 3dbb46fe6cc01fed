"""CTSim's `geom` struct (parameters/loadProjectionGeometryCT.m:14-55) as a Python object.

Fields keep the reference's names so that code written against the MATLAB struct
reads the same: reconSize, reconSpacing, reconOffset, detSize, detSpacing, detOffset,
SAD, ADD, SDD, noViews, betas, couchZ, flatPanel, map, shortScan.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import CtsimGeom


@dataclass
class Geometry:
    reconSize: tuple
    reconSpacing: tuple
    reconOffset: tuple
    detSize: tuple
    detSpacing: tuple
    detOffset: tuple
    SAD: float
    ADD: float
    betas: np.ndarray
    flatPanel: int = 1
    couchZ: np.ndarray | None = None
    map: np.ndarray | None = None
    shortScan: int = 0
    SDD: float = field(default=None)

    def __post_init__(self):
        self.reconSize = tuple(int(x) for x in np.atleast_1d(self.reconSize))
        self.reconSpacing = tuple(float(x) for x in np.atleast_1d(self.reconSpacing))
        self.reconOffset = tuple(float(x) for x in np.atleast_1d(self.reconOffset))
        self.detSize = tuple(int(x) for x in np.atleast_1d(self.detSize))
        self.detSpacing = tuple(float(x) for x in np.atleast_1d(self.detSpacing))
        self.detOffset = tuple(float(x) for x in np.atleast_1d(self.detOffset))
        if self.SDD is None:
            self.SDD = self.SAD + self.ADD  # loadProjectionGeometryCT.m:19
        self.betas = np.ascontiguousarray(self.betas, np.float64).ravel()
        if self.couchZ is None:
            self.couchZ = np.zeros(self.noViews, np.float32)
        self.couchZ = np.ascontiguousarray(self.couchZ, np.float32).ravel()

    @property
    def noViews(self) -> int:
        return int(self.betas.size)

    @property
    def ndim(self) -> int:
        """2 for fan beam (2D image, 1D detector), 3 for cone beam (geom/shell/forwardProjectMex.m:50,64)."""
        if len(self.reconSize) == 2 and len(self.detSize) == 1:
            return 2
        if len(self.reconSize) == 3 and len(self.detSize) == 2:
            return 3
        raise ValueError("Wrong geometry dimension!")

    def img_shape(self):
        return self.reconSize

    def sino_shape(self, nviews=None):
        n = self.noViews if nviews is None else nviews
        if self.ndim == 2:
            return (self.detSize[0], n)
        return (self.detSize[1], self.detSize[0], n)

    def default_map(self) -> np.ndarray:
        """geom.map default of geom/shell/forwardProjectMex.m:27-29: all columns active."""
        return np.ones(self.reconSize[:2], dtype=bool)

    def to_c(self) -> CtsimGeom:
        g = CtsimGeom()
        n, d, o = self.reconSize, self.reconSpacing, self.reconOffset
        g.nx, g.ny = n[0], n[1]
        g.dx, g.dy = d[0], d[1]
        g.ox, g.oy = o[0], o[1]
        g.nu = self.detSize[0]
        g.du = self.detSpacing[0]
        g.ou = self.detOffset[0]
        if self.ndim == 3:
            g.nz, g.dz, g.oz = n[2], d[2], o[2]
            g.nv, g.dv, g.ov = self.detSize[1], self.detSpacing[1], self.detOffset[1]
        else:
            g.nz, g.dz, g.oz = 1, 1.0, 0.0
            g.nv, g.dv, g.ov = 1, 1.0, 0.0
        g.sad, g.add, g.sdd = self.SAD, self.ADD, self.SDD
        g.detector_type = int(self.flatPanel)
        return g


def circular_betas(n: int, convention: str = "ctsim") -> np.ndarray:
    """View angles of a full circular scan.  'ctsim' is loadProjectionGeometryCT.m:33,
    ((1:N) - N/2) * 2*pi/N; 'zero' starts at 0: 2*pi*k/N."""
    k = np.arange(1, n + 1, dtype=np.float64)
    if convention == "ctsim":
        return (k - n / 2.0) * 2.0 * math.pi / n
    return (k - 1) * 2.0 * math.pi / n


def fov_map(geom: Geometry) -> np.ndarray:
    """FOV disk map of loadProjectionGeometryCT.m:42-53."""
    nx, ny = geom.reconSize[0], geom.reconSize[1]
    if geom.flatPanel:
        half = math.atan((geom.detSize[0] * geom.detSpacing[0]) / 2 / geom.SDD)
    else:
        half = (geom.detSize[0] * geom.detSpacing[0]) / 2 / geom.SDD
    fov = 2 * math.sin(half) * geom.SAD
    x = np.arange(nx) - (nx - 1) / 2
    y = np.arange(ny) - (ny - 1) / 2
    xx, yy = np.meshgrid(y, x)  # MATLAB meshgrid(y, x) -> [nx ny]
    return np.sqrt(xx ** 2 + yy ** 2) < max(max(geom.reconSize), fov / geom.reconSpacing[0]) / 2
