"""CPU tests of the geom/shell host logic: subset selection, orderedSubset and the
geometry struct (geom/shell/forwardProjectMex.m:31-39, penalty/orderedSubset.m)."""
import numpy as np

from ctsim_amd.geometry import Geometry, circular_betas, fov_map
from ctsim_amd.projector import Projector
from ctsim_amd.shell import ordered_subset, subset_views


def matlab_subset(n, M, i):
    # MATLAB: betas(1+mod(i,M):M:end), 1-based
    start = 1 + (i % M)
    return np.arange(start, n + 1, M) - 1


def test_subset_views_match_matlab_indexing():
    for n in (10, 41, 984):
        for M in (1, 2, 3, 24):
            for i in range(0, M + 1):
                want = np.arange(n) if M == 1 else matlab_subset(n, M, i)
                np.testing.assert_array_equal(subset_views(n, M, i), want)


def test_subsets_partition_the_views():
    n, M = 984, 24
    allv = np.sort(np.concatenate([subset_views(n, M, i) for i in range(1, M + 1)]))
    np.testing.assert_array_equal(allv, np.arange(n))
    assert all(len(subset_views(n, M, i)) == 41 for i in range(1, M + 1))


def test_projector_subset_tuple_matches_shell():
    for n, M in ((984, 24), (720, 7), (10, 1)):
        for i in range(1, M + 1):
            v0, vs, ns = Projector.subset(n, M, i)
            np.testing.assert_array_equal(v0 + vs * np.arange(ns), subset_views(n, M, i))


def test_ordered_subset_slices_last_axis():
    y2 = np.arange(5 * 12).reshape(5, 12)
    np.testing.assert_array_equal(ordered_subset(y2, 4, 1), y2[:, 1::4])
    np.testing.assert_array_equal(ordered_subset(y2, 4, 4), y2[:, 0::4])
    y3 = np.arange(2 * 3 * 8).reshape(2, 3, 8)
    np.testing.assert_array_equal(ordered_subset(y3, 3, 2), y3[:, :, 2::3])
    assert ordered_subset(y3, 1, 5) is y3


def test_geometry_struct_and_dimension_dispatch():
    g = Geometry(reconSize=(64, 64, 16), reconSpacing=(0.5, 0.5, 1.0), reconOffset=(0, 0, 0), detSize=(128, 32),
                 detSpacing=(0.4, 0.8), detOffset=(0, 0), SAD=1000, ADD=500, betas=circular_betas(90))
    assert g.SDD == 1500 and g.ndim == 3 and g.noViews == 90
    assert g.sino_shape() == (32, 128, 90)
    c = g.to_c()
    assert (c.nx, c.nz, c.nu, c.nv, c.detector_type) == (64, 16, 128, 32, 1)
    g2 = Geometry(reconSize=(32, 32), reconSpacing=(1, 1), reconOffset=(0, 0), detSize=(64,), detSpacing=(1,),
                  detOffset=(0,), SAD=500, ADD=300, betas=circular_betas(10))
    assert g2.ndim == 2 and g2.sino_shape() == (64, 10)
    assert g2.to_c().nz == 1 and g2.to_c().nv == 1


def test_circular_betas_ctsim_convention():
    b = circular_betas(8)
    np.testing.assert_allclose(b, (np.arange(1, 9) - 4) * 2 * np.pi / 8)
    assert np.allclose(np.diff(circular_betas(8, "zero")), 2 * np.pi / 8)


def test_fov_map_is_a_centred_disk():
    g = Geometry(reconSize=(33, 33), reconSpacing=(1, 1), reconOffset=(0, 0), detSize=(16,), detSpacing=(1,),
                 detOffset=(0,), SAD=500, ADD=300, betas=circular_betas(4))
    m = fov_map(g)
    assert m[16, 16] and not m[0, 0] and m.shape == (33, 33)
    assert (m == m.T).all()
