"""Writes tests/golden/material_tables.npz: the NIST attenuation tables of liquid water and
aluminium (a bowtie material) as the reference ships them (physicsdata/materials/, columns
energy keV, mu/rho and mu_en/rho in cm^2/g, plus the density), read with
ctsim_amd.beamhardening.read_material_attenuation.  Run in the container that holds the
reference tree:

    python tests/golden/make_material_tables.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from ctsim_amd.beamhardening import read_material_attenuation  # noqa: E402

MATERIALS = "/root/reference/physicsdata/materials"

if __name__ == "__main__":
    out = {}
    for name, key in (("Water_Liquid", "water"), ("Aluminum", "aluminum")):
        E, mu = read_material_attenuation(MATERIALS, name)
        out[f"{key}_E"] = E
        out[f"{key}_mu"] = mu
    np.savez(os.path.join(HERE, "material_tables.npz"), **out)
    print({k: v.shape for k, v in out.items()})
