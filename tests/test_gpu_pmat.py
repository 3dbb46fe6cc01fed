"""GPU parity of the projection-matrix (PMat / C-arm) operators against the oracle
(oracle/pmat_oracle.cpp, restating geom/pcbct): FP32 relative L2 <= 1e-5 for every op
string of the gateway, the matched-pair adjoint mismatch < 1e-6, on a circular orbit
written as matrices and on a C-arm-like orbit (a short arc with per-view source distance,
in-plane tilt and detector shifts folded into the matrices).  Calls go through the C ABI
(ctsim_cbct_geom_pmat_mex) via the Python mirror of the gateway."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def carm_geometry(O, n, nuv, duv, nviews, seed):
    rng = np.random.default_rng(seed)
    betas = np.linspace(-0.6, 2.2, nviews) + rng.uniform(-0.01, 0.01, nviews)
    sad = 780.0 + rng.uniform(-15, 15, nviews)
    sdd = 1180.0 + rng.uniform(-10, 10, nviews)
    P = np.zeros((3, 4, nviews))
    cams = np.zeros((3, nviews))
    for k in range(nviews):
        Pk, ck, _, _ = O.circular_pmat([betas[k]], sad[k], sdd[k], nuv, duv,
                                       (rng.uniform(-2, 2), rng.uniform(-1.5, 1.5)))
        # a small rotation of the detector plane about its normal (u/v mixing)
        a = rng.uniform(-0.02, 0.02)
        R = np.array([[np.cos(a), -np.sin(a), 0.0], [np.sin(a), np.cos(a), 0.0], [0.0, 0.0, 1.0]])
        c = np.array([(nuv[0] - 1) / 2.0, (nuv[1] - 1) / 2.0])
        T1 = np.array([[1, 0, -c[0]], [0, 1, -c[1]], [0, 0, 1.0]])
        T2 = np.array([[1, 0, c[0]], [0, 1, c[1]], [0, 0, 1.0]])
        P[:, :, k] = T2 @ R @ T1 @ Pk[:, :, 0]
        cams[:, k] = ck[:, 0]
    ouv = np.stack([rng.uniform(-1, 1, nviews), rng.uniform(-1, 1, nviews)])
    return P, cams, sdd, ouv


@pytest.mark.parametrize("kind", ["circular", "carm"])
def test_pmat_ops_match_oracle(oracle_mod, kind):
    O = oracle_mod
    from ctsim_amd import mex
    n, d, o = (40, 36, 20), (0.7, 0.7, 0.9), (0.5, -0.25, 0.5)
    nuv, duv = (72, 36), (0.8, 0.9)
    nviews = 14
    if kind == "circular":
        P, cams, sdds, ouv = O.circular_pmat(np.linspace(0, 2 * np.pi, nviews, endpoint=False) + 0.05, 1000.0,
                                             1500.0, nuv, duv, (0.5, -0.25))
    else:
        P, cams, sdds, ouv = carm_geometry(O, n, nuv, duv, nviews, 3)
    m = np.ones(n[:2], bool)
    m[:3, :] = False
    m[:, -2:] = False
    g = O.GP(n, d, o, nuv, duv[0], duv[1], P, cams, sdds, ouv, map_=m)
    rng = np.random.default_rng(11)
    vol = np.asfortranarray(rng.uniform(0, 1, n).astype(np.float32))
    vol[rng.uniform(size=n) < 0.1] = 0.0  # the forward skips zero voxels
    sino = np.asfortranarray(rng.uniform(0, 1, g.sino_shape()).astype(np.float32))
    args = (n, d, o, nuv, duv, nviews, g.pmat, g.cameras, g.sdds, g.ouv)
    print(f"pmat {kind}: n={n} nuv={nuv} views={nviews}", flush=True)
    got = {}
    for op in ("proj,dd", "proj,tf", "back,dd", "back,tf", "back,pd", "back,hpd"):
        data = vol if op.startswith("proj") else sino
        got[op] = mex.cbct_geom_pmat_mex(op, *args, data, m)
        want = O.pmat(op, g, data)
        e = rel_l2(got[op], want)
        print(f"  {op}: rel-L2 vs FP32 oracle {e:.2e}", flush=True)
        assert e <= 1e-5, (op, e)
        assert np.linalg.norm(want) > 0
    ax, aty = got["proj,dd"].astype(np.float64), got["back,dd"].astype(np.float64)
    mis = abs(np.vdot(ax, sino) - np.vdot(vol, aty)) / (np.linalg.norm(ax) * np.linalg.norm(sino))
    print(f"  adjoint mismatch {mis:.2e}", flush=True)
    assert mis < 1e-6
    assert np.array_equal(got["back,dd"], got["back,tf"])
    with pytest.raises(Exception, match="ray driven method is not supported"):
        mex.cbct_geom_pmat_mex("proj,rd", *args, vol, m)


@pytest.mark.parametrize("case", ["tall", "fine", "steep", "thin"])
def test_pmat_forward_tiles_chunks_and_fallback(oracle_mod, case):
    """The forward's tiled form against the oracle where the easy case does not reach:
    "tall" has 150 slices (three slice chunks, the last one partial) and a detector that
    cuts the top and bottom slices; "fine" has 0.05 mm cells, so one voxel covers about 30
    of them (the footprint truncated at 16 u cells and 10 v rows, as the reference does) and
    a tile's detector box exceeds the LDS patch, which sends every cell to global atomics;
    "steep" has 2 mm slices over 0.5 mm rows (six rows a voxel, past the in-patch row walk,
    so LDS atomics inside the patch); "thin" has 0.3 mm slices over 1 mm rows (about two
    slices a row, so neighbouring lanes share a first row and the plain adds are off)."""
    O = oracle_mod
    from ctsim_amd import mex
    if case == "tall":
        n, d, o, nuv, duv, nviews = (24, 20, 150), (1.0, 1.0, 0.9), (0.3, -0.2, 0.0), (48, 180), (1.0, 1.0), 6
    elif case == "fine":
        n, d, o, nuv, duv, nviews = (20, 18, 10), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), (400, 200), (0.05, 0.05), 4
    elif case == "steep":
        n, d, o, nuv, duv, nviews = (12, 10, 30), (0.5, 0.5, 2.0), (0.0, 0.0, 0.0), (64, 200), (0.5, 0.5), 5
    else:
        n, d, o, nuv, duv, nviews = (20, 18, 70), (1.0, 1.0, 0.3), (0.0, 0.0, 0.0), (48, 40), (1.0, 1.0), 5
    P, cams, sdds, ouv = O.circular_pmat(np.linspace(0, 2 * np.pi, nviews, endpoint=False) + 0.3, 1000.0, 1500.0,
                                         nuv, duv)
    m = np.ones(n[:2], bool)
    g = O.GP(n, d, o, nuv, duv[0], duv[1], P, cams, sdds, ouv, map_=m)
    rng = np.random.default_rng(5)
    vol = np.asfortranarray(rng.uniform(0, 1, n).astype(np.float32))
    args = (n, d, o, nuv, duv, nviews, g.pmat, g.cameras, g.sdds, g.ouv)
    print(f"pmat forward {case}: n={n} nuv={nuv} views={nviews}", flush=True)
    got = mex.cbct_geom_pmat_mex("proj,dd", *args, vol, m)
    want = O.pmat("proj,dd", g, vol)
    e = rel_l2(got, want)
    print(f"  proj,dd: rel-L2 vs FP32 oracle {e:.2e}", flush=True)
    assert np.linalg.norm(want) > 0
    assert e <= 1e-5
