"""Water beam-hardening correction of the raw-data path (sinogram/log/processCTRawData.m:40-43).

The host part restates the reference's polynomial fit, once per spectrum and bowtie
channel; the correction itself runs on the device (ctsim_water_correction):

  read_material_attenuation   phantom/readMaterialAttenuation.m:1-60 (NIST ASCII tables:
                              density_<name>.txt, attenuation_<name>.txt)
  spectral_attenuation        spectra/spectralAttenuations.m with
                              artifact/beamhardening/interp1geom.m (log-log interpolation)
  resulting_photons           spectra/computeResultingPhotons.m (mm thicknesses, 1/cm mu)
  water_coefficients          artifact/beamhardening/beamHardeningCorrectionPolynomialCoefficients.m
  water_coefficient_table     artifact/beamhardening/beamHardeningWarterCorrection.m:21-48
  water_correction            the polynomial on a device sinogram, per detector channel

A material table is read from a directory the caller names (CTSim ships them in
physicsdata/materials/); nothing here carries one.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import F32, call
from .device import DeviceArray


def read_material_attenuation(materials_dir: str, name: str):
    """(E keV, mu 1/cm) of a NIST material table: E from column 1 (absorption-edge
    duplicates pulled apart by a factor (1 + 2 eps) each way), mu = (mu/rho) * rho."""
    with open(os.path.join(materials_dir, f"density_{name}.txt")) as f:
        rho = float(f.read().split()[0])
    data = np.loadtxt(os.path.join(materials_dir, f"attenuation_{name}.txt"), dtype=np.float64, ndmin=2)
    E = data[:, 0].copy()
    eps = np.finfo(np.float64).eps
    for i in range(len(E) - 1):
        if E[i] == E[i + 1]:
            E[i] = E[i] / (1 + 2 * eps)
            E[i + 1] = E[i + 1] * (1 + 2 * eps)
    return E, data[:, 1] * rho


def interp_geometric(x, y, xi):
    """exp(interp1(log x, log y, log xi)) with MATLAB's default NaN outside [x_1, x_n]."""
    x, y, xi = (np.asarray(a, np.float64) for a in (x, y, xi))
    lx, lxi = np.log(x), np.log(xi)
    out = np.exp(np.interp(lxi, lx, np.log(y)))
    return np.where((lxi < lx[0]) | (lxi > lx[-1]), np.nan, out)


def spectral_attenuation(energies, E, mu):
    """mu at each spectrum energy (keV), from the table (E, mu)."""
    return interp_geometric(E, mu, np.atleast_1d(energies))


def resulting_photons(photons, energies, layers=()):
    """photons * exp(-sum_m mu_m(E) t_m / 10) for layers of (E_table, mu_table, thickness mm)."""
    e = np.atleast_1d(np.asarray(energies, np.float64))
    att = np.zeros_like(e)
    for E, mu, t in layers:
        att += spectral_attenuation(e, E, mu) * (float(t) / 10.0)
    return np.asarray(photons, np.float64) * np.exp(-att)


def water_coefficients(energies, photons_per_bin, E_water, mu_water, energy_eff, order=3, max_thickness=50.0):
    """Least-squares coefficients c_1..c_order (no constant term) that map the polychromatic
    water line integral to the monochromatic one at energy_eff, over 256 water thicknesses
    0 .. max_thickness cm (beamHardeningCorrectionPolynomialCoefficients.m:1-43; the normal
    equations solved as the reference solves them, (A'A) \\ (A'b))."""
    p = np.asarray(photons_per_bin, np.float64)
    total = np.sum(p)
    e = np.atleast_1d(np.asarray(energies, np.float64))
    mu_poly = spectral_attenuation(e, E_water, mu_water)
    # physicsdata/materialAttenuation.m:35-45: the effective attenuation is interpolated
    # over the bin energies' own attenuations (one bin: that bin's, whatever energy_eff)
    mu_mono = float(mu_poly[0]) if e.size == 1 else float(interp_geometric(e, mu_poly, [energy_eff])[0])
    lengths = np.linspace(0.0, max_thickness, 256)
    i_poly = np.array([np.sum(p * np.exp(-length * mu_poly)) for length in lengths])
    i_mono = total * np.exp(-lengths * mu_mono)
    lp = -np.log(i_poly / total)
    lm = -np.log(i_mono / total)
    A = np.stack([lp ** (i + 1) for i in range(order)], axis=1)
    return np.linalg.solve(A.T @ A, A.T @ lm)


def water_coefficient_table(energies, photons_per_bin_original, E_water, mu_water, energy_eff, dqe=1.0, gain=1.0,
                            energy_integrating=True, bowtie=None, order=3, max_thickness=50.0,
                            multi_channels=True):
    """Coefficients per detector channel (multi_channels, the form processCTRawData uses:
    one fit per bowtie thickness) or one set for the mean bowtie thickness
    (beamHardeningWarterCorrection.m:21-48).  bowtie = (E_table, mu_table, thicknesses mm
    per channel) or None.  Returns [nchannels, order] (multi) or [1, order]."""
    e = np.atleast_1d(np.asarray(energies, np.float64))
    ph = np.asarray(photons_per_bin_original, np.float64) * dqe * gain
    if energy_integrating:
        ph = ph * e
    if bowtie is None:
        return water_coefficients(e, ph, E_water, mu_water, energy_eff, order, max_thickness)[None, :]
    Eb, mub, thick = bowtie
    thick = np.atleast_1d(np.asarray(thick, np.float64))
    if not multi_channels:
        pc = resulting_photons(ph, e, [(Eb, mub, float(np.mean(thick)))])
        return water_coefficients(e, pc, E_water, mu_water, energy_eff, order, max_thickness)[None, :]
    return np.stack([water_coefficients(e, resulting_photons(ph, e, [(Eb, mub, t)]), E_water, mu_water, energy_eff,
                                        order, max_thickness) for t in thick])


def water_correction(sino: DeviceArray, nv: int, nu: int, nviews: int, coeffs: np.ndarray,
                     out: DeviceArray | None = None, stream=None) -> DeviceArray:
    """Apply the polynomial on the device: coeffs [nu, order] (per channel) or [1, order]."""
    c = np.ascontiguousarray(np.atleast_2d(coeffs), np.float32)
    per_channel = c.shape[0] > 1
    if per_channel and c.shape[0] != nu:
        raise ValueError("one coefficient row per detector channel (nu rows) or a single row")
    if sino.count != nv * nu * nviews:
        raise ValueError("sinogram size does not match nv * nu * nviews")
    d_c = DeviceArray.from_numpy(c.ravel(), F32, stream=stream)
    out = out if out is not None else DeviceArray(sino.count, sino.dtype, shape=sino.shape)
    call("ctsim_water_correction", sino.vp, sino.dtype, out.vp, out.dtype, int(nv), int(nu), int(nviews), d_c.vp,
         int(c.shape[1]), int(per_channel), C.c_void_p(stream.handle if stream else 0))
    if stream is not None:
        stream.synchronize()  # the coefficient buffer is freed on return
    return out
