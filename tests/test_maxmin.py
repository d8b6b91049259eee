"""maxmin_order (vecchia.rs:144-215) of the library — a host computation,
so these run without a GPU — against the oracle's restatement of the
reference's lazy traversal: identical orders on random, clustered and
lattice (tied-distance) point sets in 2-D and 3-D, plus the degenerate
sizes and the rejections."""
import numpy as np
import pytest

import _cases as cs
import geostat_hip as gh


@pytest.mark.parametrize("shape,D,n", [("uniform", 2, 20000), ("uniform", 3, 8000),
                                       ("clusters", 2, 6000), ("lattice", 2, 3000),
                                       ("lattice", 3, 2744)])
def test_maxmin_matches_oracle(oracle, shape, D, n):
    r = np.random.default_rng(n + D)
    if shape == "uniform":
        c = r.uniform(0.0, 1.0, size=(n, D))
    elif shape == "clusters":
        centres = r.uniform(0.0, 1.0, size=(6, D))
        c = np.unique(centres[r.integers(6, size=n)] + 0.02 * r.standard_normal((n, D)), axis=0)
    else:
        k = round(n ** (1.0 / D))
        axes = np.meshgrid(*[np.arange(k, dtype=float)] * D, indexing="ij")
        c = np.column_stack([a.ravel() for a in axes])
        r.shuffle(c)
    got = gh.maxmin_order(c)
    want = oracle.maxmin_order(c)
    assert np.array_equal(got, np.asarray(want, dtype=np.int64))


def test_maxmin_small_sizes_and_rejections(oracle):
    assert gh.maxmin_order(np.zeros((0, 2))).size == 0
    assert list(gh.maxmin_order([[3.0, 4.0]])) == [0]
    c, _ = cs.sample_data()
    assert list(gh.maxmin_order(c)) == list(oracle.maxmin_order(c))
    assert list(gh.maxmin_order(c[:2])) == list(oracle.maxmin_order(c[:2]))
    with pytest.raises(gh.InvalidParameter):
        gh.maxmin_order([[0.0, np.nan], [1.0, 1.0]])
    with pytest.raises(gh.DimensionMismatch):
        gh.maxmin_order(np.zeros((4, 4)))
