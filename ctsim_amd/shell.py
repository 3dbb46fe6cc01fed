"""Python counterpart of CTSim's MATLAB operator boundary (geom/shell/).

forward_project_mex / back_project_mex reproduce geom/shell/forwardProjectMex.m:1-93 and
geom/shell/backProjectMex.m:1-195: default map, ordered-subset view selection
``betas(1+mod(i,M):M:end)``, 2D / 3D dispatch to the MEX gateways, and the factor M on
subset backprojections.  load_projectors is geom/shell/loadPojectors.m:1-12 and
ordered_subset is penalty/orderedSubset.m:1-9.  Arrays are numpy (MATLAB order).
"""
from __future__ import annotations

import numpy as np

from . import mex
from .geometry import Geometry


def subset_views(n: int, M: int, i: int) -> np.ndarray:
    """0-based view indices of subset i of M: MATLAB 1+mod(i,M):M:end."""
    if M == 1:
        return np.arange(n)
    return np.arange(i % M, n, M)


def _select(geom: Geometry, M: int, i: int):
    if M == 1:
        return geom.betas, geom.noViews, geom.couchZ
    idx = subset_views(geom.noViews, M, i)
    return geom.betas[idx], idx.size, geom.couchZ[idx]


def forward_project_mex(img, geom: Geometry, M: int = 1, i: int = 0, proj_type: str = "proj,dd", map=None):
    if map is not None:
        geom.map = map
    if geom.map is None:
        geom.map = geom.default_map()
    betas, nviews, couchz = _select(geom, M, i)
    img = np.asarray(img, np.float32)
    if geom.ndim == 2:
        return mex.fbct_geom_mex(proj_type, np.int32(geom.reconSize), np.float32(geom.reconSpacing),
                                 np.float32(geom.reconOffset), np.int32(geom.detSize), np.float32(geom.detSpacing),
                                 np.float32(geom.detOffset), np.float32(geom.SAD), np.float32(geom.ADD),
                                 np.float32(geom.SDD), np.int32(nviews), np.float64(betas),
                                 np.int32(geom.flatPanel), img)
    return mex.cbct_geom_mex(proj_type, np.int32(geom.reconSize), np.float32(geom.reconSpacing),
                             np.float32(geom.reconOffset), np.int32(geom.detSize), np.float32(geom.detSpacing),
                             np.float32(geom.detOffset), np.float32(geom.SAD), np.float32(geom.ADD),
                             np.float32(geom.SDD), np.int32(nviews), np.float64(betas), np.float32(couchz),
                             np.int32(geom.flatPanel), img, np.asarray(geom.map, bool))


def back_project_mex(sino, geom: Geometry, M: int = 1, i: int = 0, proj_type: str = "back,dd", map=None):
    if map is not None:
        geom.map = map
    if geom.map is None:
        geom.map = geom.default_map()
    betas, nviews, couchz = _select(geom, M, i)
    sino = np.asarray(sino, np.float32)
    if geom.ndim == 2:
        img = mex.fbct_geom_mex(proj_type, np.int32(geom.reconSize), np.float32(geom.reconSpacing),
                                np.float32(geom.reconOffset), np.int32(geom.detSize), np.float32(geom.detSpacing),
                                np.float32(geom.detOffset), np.float32(geom.SAD), np.float32(geom.ADD),
                                np.float32(geom.SDD), np.int32(nviews), np.float64(betas),
                                np.int32(geom.flatPanel), sino)
    else:
        img = mex.cbct_geom_mex(proj_type, np.int32(geom.reconSize), np.float32(geom.reconSpacing),
                                np.float32(geom.reconOffset), np.int32(geom.detSize), np.float32(geom.detSpacing),
                                np.float32(geom.detOffset), np.float32(geom.SAD), np.float32(geom.ADD),
                                np.float32(geom.SDD), np.int32(nviews), np.float64(betas), np.float32(couchz),
                                np.int32(geom.flatPanel), sino, np.asarray(geom.map, bool))
    if M != 1:
        img = (np.float32(M) * img).astype(np.float32)
    return img


def ordered_subset(x, M: int, isub: int):
    """penalty/orderedSubset.m: the views of subset isub along the last axis."""
    if M == 1:
        return x
    x = np.asarray(x)
    if x.ndim == 2:
        return x[:, isub % M::M]
    if x.ndim == 3:
        return x[:, :, isub % M::M]
    raise ValueError("orderedSubset expects a 2D or 3D sinogram")


def load_projectors(geom: Geometry, numos: int = 1):
    """geom/shell/loadPojectors.m: separable-footprint operators A, At, Aos, Atos, Os."""
    A = lambda x: forward_project_mex(x, geom, 1, 0, "proj,tf")  # noqa: E731
    At = lambda x: back_project_mex(x, geom, 1, 0, "back,tf")  # noqa: E731
    Aos = lambda x, isub: forward_project_mex(x, geom, numos, isub, "proj,tf")  # noqa: E731
    Atos = lambda x, isub: back_project_mex(x, geom, numos, isub, "back,tf")  # noqa: E731
    Os = lambda x, isub: ordered_subset(x, numos, isub)  # noqa: E731
    return A, At, Aos, Atos, Os
