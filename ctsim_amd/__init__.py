"""MI355X-native (gfx950 HIP) implementation of CTSim's projector / backprojector path.

Layers:
  mex        - the MEX gateways' argument lists (fbct_geom_mex, cbct_geom_mex, ...)
  shell      - geom/shell counterparts (forward_project_mex, back_project_mex, ...)
  projector  - device-resident plans for loops that keep data in HBM
  pwls, fdk  - the reconstruction loops that call the operators
All compute goes through libctsim_hip.so (C ABI in include/ctsim_amd.h).
"""
from ._lib import BF16, F16, F32, F64, CtsimError, lib  # noqa: F401
from .geometry import Geometry, circular_betas, fov_map  # noqa: F401

__all__ = ["Geometry", "circular_betas", "fov_map", "CtsimError", "lib", "F32", "F16", "BF16", "F64"]
