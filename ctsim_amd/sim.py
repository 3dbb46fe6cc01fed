"""Raw-data side of the simulation (sinogram/log/processCTRawData.m:1-43) on the GPU.

source_intensity() is the expected unattenuated count of the spectrum, as the two log
functions compute it on the host; log_attenuation() runs the device log with the
tube-current profile, the bowtie flat-field ratio and the clamp of the chosen noise
model.  process_ct_raw_data() chains it with the water beam-hardening correction
(beamhardening.water_correction, coefficients per detector channel from
beamhardening.water_coefficient_table), as processCTRawData.m:37-40 does.  The compound
model takes the scintillator's attenuation at the bin energies from the caller.
"""
from __future__ import annotations

import numpy as np

from ._lib import call
from .device import DeviceArray, dtype_code

SIMPLE_LOW_COUNT = 10.0     # computeSimplePoissonSinogramAttunation.m:14
COMPOUND_LOW_COUNT = 100.0  # computeCompoundPoissonSinogramAttunation.m:14
SCINTILLATOR_CM = 1.5       # hDetectorEnergyResponseFunction.m:3
GEOMETRY_LOSS = 0.90        # hDetectorEnergyResponseFunction.m:5


def source_intensity(photons_per_bin, energies, dqe=1.0, gain=1.0, energy_integrating=True,
                     compound=False, scintillator_mu=None, photons_total=None) -> float:
    """Expected unattenuated count of one detector cell.

    Simple model (computeSimplePoissonSinogramAttunation.m:18-37): one energy bin gives
    dqe * photons_total; several give dqe * sum(photons * energies) * gain (energy
    integrating) or dqe * sum(photons) * gain (photon counting).  Compound model
    (computeCompoundPoissonSinogramAttunation.m:16-19 with hDetectorEnergyResponseFunction.m):
    sum(0.9 * photons * (1 - exp(-mu * 1.5)) * energies * gain), mu being the scintillator's
    linear attenuation (1/cm) at each bin energy."""
    e = np.atleast_1d(np.asarray(energies, np.float64))
    p = np.atleast_1d(np.asarray(photons_per_bin, np.float64))
    if compound:
        if scintillator_mu is None:
            raise ValueError("the compound model needs the scintillator attenuation at each bin energy")
        mu = np.atleast_1d(np.asarray(scintillator_mu, np.float64))
        detected = GEOMETRY_LOSS * p * (1.0 - np.exp(-mu * SCINTILLATOR_CM))
        return float(np.sum(detected * e * gain))
    if e.size == 1:
        total = float(p[0]) if photons_total is None else float(photons_total)
        return dqe * total
    if energy_integrating:
        return dqe * float(np.sum(p * e)) * gain
    return dqe * float(np.sum(p)) * gain


def log_attenuation(counts: DeviceArray, ncell: int, source_intensity: float, compound: bool = False,
                    low_count: float | None = None, flat_field: DeviceArray | None = None,
                    tube_profile: DeviceArray | None = None, out_dtype: str = "f32", stream=None) -> DeviceArray:
    """-log of the normalised counts, [nviews][ncell] (processCTRawData.m:14-35 without the
    water correction).  counts is FP32; flat_field holds ncell ratios, tube_profile one
    value per view."""
    if counts.dtype != dtype_code("f32"):
        raise ValueError("counts must be FP32")
    if low_count is None:
        low_count = COMPOUND_LOW_COUNT if compound else SIMPLE_LOW_COUNT
    if flat_field is not None and flat_field.count != ncell:
        raise ValueError("flat_field must hold one ratio per detector cell")
    if tube_profile is not None and tube_profile.count * ncell != counts.count:
        raise ValueError("tube_profile must hold one value per view")
    y = DeviceArray(counts.count, out_dtype, shape=counts.shape)
    call("ctsim_log_attenuation_ex", counts.vp, counts.count, int(ncell), float(source_intensity),
         float(low_count), int(bool(compound)), flat_field.vp if flat_field is not None else None,
         tube_profile.vp if tube_profile is not None else None, y.vp, y.dtype,
         stream.vp if stream is not None else None)
    return y


def process_ct_raw_data(counts: DeviceArray, nv: int, nu: int, source_intensity: float, compound: bool = False,
                        flat_field: DeviceArray | None = None, tube_profile: DeviceArray | None = None,
                        water_coeffs: np.ndarray | None = None, out_dtype: str = "f32", stream=None) -> DeviceArray:
    """processCTRawData.m:1-43 on a device sinogram [nviews][nu][nv] (nv = 1 for fan beam):
    the tube-current division and -log (log_attenuation), then, when water_coeffs is given
    ([nu, order] per channel, or [1, order]), the water polynomial in place."""
    from .beamhardening import water_correction
    y = log_attenuation(counts, nv * nu, source_intensity, compound, None, flat_field, tube_profile, out_dtype,
                        stream)
    if water_coeffs is not None:
        water_correction(y, nv, nu, counts.count // (nv * nu), water_coeffs, out=y, stream=stream)
    return y
