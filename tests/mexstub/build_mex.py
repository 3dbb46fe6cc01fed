"""Compile the mexFunction shims of ctsim_amd/csrc/mex/ against the stand-in mex.h.

Outputs (git-ignored, next to this file in _lib/): libmexstub.so (the array model of
mexstub.c) and one shared object per shim, each exporting mexFunction and linked
against libmexstub.so and ctsim_amd/libctsim_hip.so.  gcc -Wall -Werror as C11: a shim
whose calls drift from include/ctsim_amd.h or from MATLAB's mx* signatures fails here.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
MEX_SRC = os.path.join(ROOT, "ctsim_amd", "csrc", "mex")
LIB_DIR = os.path.join(HERE, "_lib")
SHIMS = ["fbct_geom_mex", "cbct_geom_mex", "cbct_dd_cuda", "cbct_sf_cuda", "cbct_back_pd_cuda", "cbct_geom_pmat_mex"]
CFLAGS = ["-std=c11", "-O1", "-fPIC", "-Wall", "-Wextra", "-Werror", "-Wno-unused-parameter",
          "-Wno-unused-function"]


def _run(cmd, out):
    # written under a process-unique name and renamed: concurrent test workers never load a
    # half-written object
    tmp = f"{out}.{os.getpid()}.tmp"
    r = subprocess.run([*cmd, "-o", tmp], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, out)


def build() -> dict:
    """Returns {shim name: path of its shared object}."""
    os.makedirs(LIB_DIR, exist_ok=True)
    stub = os.path.join(LIB_DIR, "libmexstub.so")
    _run(["gcc", *CFLAGS, "-shared", "-I", HERE, os.path.join(HERE, "mexstub.c")], stub)
    lib_dir = os.path.join(ROOT, "ctsim_amd")
    out = {}
    for name in SHIMS:
        so = os.path.join(LIB_DIR, f"{name}.so")
        _run(["gcc", *CFLAGS, "-shared", "-I", HERE, "-I", os.path.join(ROOT, "include"),
              os.path.join(MEX_SRC, f"{name}.c"), "-L", LIB_DIR, "-lmexstub", "-L", lib_dir,
              "-l:libctsim_hip.so", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/../../../ctsim_amd"], so)
        out[name] = so
    return out


if __name__ == "__main__":
    for k, v in build().items():
        print(k, v)
