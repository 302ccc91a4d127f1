"""Host side of the synthetic workload generators (CPU): the DEM form and the ldd it yields."""
import numpy as np

from lue_amd import synthetic


def test_dem_has_no_interior_pits_and_drains_south():
    ldd = synthetic.ldd_host(20261203, 64, 300)
    assert (ldd[-1] == 5).all()
    assert (ldd[:-1] != 5).all()  # e(r,c) - e(r+1,c) >= 0.1: a lower neighbour always exists
    # no cell points north: every drop to a northern neighbour is negative
    assert not np.isin(ldd, [7, 8, 9]).any()


def test_uniform_is_seeded_and_in_range():
    a = synthetic.uniform_host(7, 1, 0.0, 5.0, (100, 100))
    b = synthetic.uniform_host(7, 1, 0.0, 5.0, (100, 100))
    c = synthetic.uniform_host(7, 2, 0.0, 5.0, (100, 100))
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= 0 and a.max() < 5
    assert abs(a.mean() - 2.5) < 0.1
