"""FP64 numpy restatement of the raw-data log.

TEST INFRASTRUCTURE ONLY (imported by tests/): the checker of ctsim_amd.sim.

  process_raw(raw, i0, profile, ff, compound, low)
      sinogram/log/processCTRawData.m:14-33 (division by the tube-current profile per
      view, then one of the two models below; the water correction is not restated)
      simple:   computeSimplePoissonSinogramAttunation.m:14-52 (clamp the counts at low,
                divide by the source intensity, then by the bowtie flat-field ratio, -log)
      compound: computeCompoundPoissonSinogramAttunation.m:14-38 (divide by the source
                intensity and the flat-field ratio, clamp the ratio at low / i0, -log)
  raw is [nviews][ncell], ff [ncell] (the flatFieldRatio in the order of the detector
  cells), profile [nviews].
"""
from __future__ import annotations

import numpy as np


def process_raw(raw, i0, profile=None, ff=None, compound=False, low=10.0):
    c = np.asarray(raw, np.float64)
    if profile is not None:
        c = c / np.asarray(profile, np.float64)[:, None]
    f = np.ones(c.shape[1]) if ff is None else np.asarray(ff, np.float64)
    if compound:
        r = (c / i0) / f[None, :]
        return -np.log(np.maximum(r, low / i0))
    r = (np.maximum(c, low) / i0) / f[None, :]
    return -np.log(r)
