"""Kernel-time workload for the operators off the benchmark path (run under a profiler;
pytest does not collect it): the projection-matrix (C-arm) forward and back projectors on
a circular orbit written as matrices, and the ray-driven forward projector, each at a
256^3 volume (1 mm), 360 views, a flat 512 x 384 detector at 1 mm, SAD 1000 / SDD 1500 mm,
through the MATLAB gateways (host arrays in and out; the profiler's kernel trace isolates
the kernels).  The matrices come from the oracle's circular-orbit builder; nothing is
checked here (tests/test_gpu_pmat.py and tests/test_gpu_parity.py hold the parity tests).

    python tests/perf_pmat_rd.py [reps]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from ctsim_amd import mex  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main(reps: int) -> None:
    n, d, o = (256, 256, 256), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    nuv, duv, nviews = (512, 384), (1.0, 1.0), 360
    betas = np.linspace(0, 2 * np.pi, nviews, endpoint=False)
    P, cams, sdds, ouv = O.circular_pmat(betas, 1000.0, 1500.0, nuv, duv)
    m = np.ones(n[:2], bool)
    g = O.GP(n, d, o, nuv, duv[0], duv[1], P, cams, sdds, ouv, map_=m)  # the gateway's single arrays
    args = (n, d, o, nuv, duv, nviews, g.pmat, g.cameras, g.sdds, g.ouv)
    rng = np.random.default_rng(0)
    vol = np.asfortranarray(rng.uniform(0, 1, n).astype(np.float32))
    for _ in range(reps):
        sino = mex.cbct_geom_pmat_mex("proj,dd", *args, vol, m)
        mex.cbct_geom_pmat_mex("back,dd", *args, sino, m)
        mex.cbct_geom_pmat_mex("back,hpd", *args, sino, m)
        mex.cbct_geom_mex("proj,rd", n, d, o, nuv, duv, (0.0, 0.0), 1000.0, 500.0, 1500.0, nviews, betas,
                          np.zeros(nviews, np.float32), 1, vol)  # couch z: single, as the gateway checks
    print("done", reps, "x pmat proj,dd / back,dd / back,hpd and proj,rd at", n, nuv, nviews, flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 2)
