"""maxmin_order (vecchia.rs:144-215) computed on the GPU (gs_maxmin_order_gpu):
the same order as the oracle's restatement of the reference's lazy traversal
on random, clustered, lattice (tied-distance) and duplicated point sets in
2-D and 3-D, with the host's finishing traversal forced at several rounds,
and at the full 10^6-point Vecchia workload the same order as the library's
host traversal (itself checked against the oracle without a GPU)."""
import os

import numpy as np
import pytest

import geostat_hip as gh
from geostat_hip import workloads

pytestmark = pytest.mark.gpu


def _points(shape, D, n, seed):
    r = np.random.default_rng(seed)
    if shape == "uniform":
        return r.uniform(0.0, 1.0, size=(n, D))
    if shape == "clusters":
        centres = r.uniform(0.0, 1.0, size=(6, D))
        return np.unique(centres[r.integers(6, size=n)] + 0.02 * r.standard_normal((n, D)), axis=0)
    if shape == "lattice":
        k = round(n ** (1.0 / D))
        axes = np.meshgrid(*[np.arange(k, dtype=float)] * D, indexing="ij")
        c = np.column_stack([a.ravel() for a in axes])
        r.shuffle(c)
        return c
    if shape == "duplicates":      # every point twice (the plan rejects these; the order does not)
        c = r.uniform(0.0, 1.0, size=(n // 2, D))
        return r.permutation(np.concatenate([c, c]))
    raise ValueError(shape)


@pytest.mark.parametrize("shape,D,n", [("uniform", 2, 20000), ("uniform", 3, 8000),
                                       ("uniform", 2, 3), ("uniform", 2, 2), ("uniform", 3, 65),
                                       ("clusters", 2, 6000), ("clusters", 3, 6000),
                                       ("lattice", 2, 3000), ("lattice", 3, 2744),
                                       ("duplicates", 2, 3000)])
def test_gpu_maxmin_matches_oracle(oracle, shape, D, n):
    c = _points(shape, D, n, n + D)
    print(f"maxmin gpu vs oracle: {shape} D={D} n={len(c)}", flush=True)
    got, info = gh.maxmin_order_gpu(c, info=True)
    want = np.asarray(oracle.maxmin_order(c), dtype=np.int64)
    assert np.array_equal(got, want), info
    assert info[2] >= 1 and info[0] >= 0
    if shape in ("uniform", "clusters") and len(c) > 1000:
        assert info[2] == len(c), info      # no hand-over to the host


@pytest.mark.parametrize("after", [0, 1, 7, 25])
def test_gpu_maxmin_host_finish(oracle, after, monkeypatch):
    # the host finishes from the device's keys after `after` rounds
    c = _points("uniform", 2, 100000, 11)
    monkeypatch.setenv("GS_MAXMIN_HOST_AFTER", str(after))
    print(f"maxmin gpu host finish after {after} rounds", flush=True)
    got, info = gh.maxmin_order_gpu(c, info=True)
    assert info[2] < len(c) and info[0] <= max(after, 1)
    assert np.array_equal(got, gh.maxmin_order(c))
    assert np.array_equal(got, np.asarray(oracle.maxmin_order(c), dtype=np.int64))


def test_gpu_maxmin_small_and_rejections():
    assert gh.maxmin_order_gpu(np.zeros((0, 2))).size == 0
    assert list(gh.maxmin_order_gpu([[3.0, 4.0]])) == [0]
    with pytest.raises(gh.InvalidParameter):
        gh.maxmin_order_gpu([[0.0, np.nan], [1.0, 1.0]])
    with pytest.raises(gh.DimensionMismatch):
        gh.maxmin_order_gpu(np.zeros((4, 4)))


def test_gpu_maxmin_full_size():
    c = workloads.vecchia_c3()
    print(f"maxmin gpu full size: n={len(c)}", flush=True)
    got, info = gh.maxmin_order_gpu(c, info=True)
    assert info[2] == len(c)          # ordered on the GPU to the end
    assert np.array_equal(got, gh.maxmin_order(c))
    again = gh.maxmin_order_gpu(c)
    assert np.array_equal(got, again)
