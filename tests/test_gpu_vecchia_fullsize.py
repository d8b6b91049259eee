"""Size-independent properties of the Vecchia plan at the full 10^6-point
workload (the oracle's serial plan build is too slow at this size):

* the order is a permutation and its first point is the one nearest to the
  centroid;
* maxmin: the distance of each ordered point to its nearest predecessor (the
  first entry of its conditioning set) never increases along the order;
* conditioning sets (point indices): min(k, m) predecessors, and, for a
  sample of positions, exactly the m nearest predecessors by brute force in
  ascending (d2, position) order;
* the likelihood batch is deterministic across calls;
* the log-likelihoods of two of the 64 parameter vectors equal the oracle's
  (the reference's serial loop, restated) on the same plan within 1e-10.
"""
import numpy as np
import pytest

import geostat_hip as gh
from geostat_hip import workloads

pytestmark = pytest.mark.gpu


def test_full_size_plan_properties():
    c = workloads.vecchia_c3()
    n, m = len(c), 30
    print(f"vecchia full size: n={n} m={m}", flush=True)
    p = gh.vecchia_plan(c, m)
    o, nb, cnt = p.arrays()
    assert np.array_equal(np.sort(o), np.arange(n))
    centre = c.mean(0)
    assert o[0] == int(np.argmin(((c - centre) ** 2).sum(1)))
    pos = np.arange(n)
    assert np.array_equal(cnt, np.minimum(pos, m))
    x = c[o]                                  # coordinates in ordered position
    inv = np.empty(n, dtype=np.int64)
    inv[o] = pos
    valid = nb >= 0
    npos = np.where(valid, inv[np.where(valid, nb, 0)], -1)   # neighbours as positions
    assert (valid.sum(1) == cnt).all() and (npos < pos[:, None])[valid].all()
    # nearest-predecessor distance never increases (maxmin)
    d_near = ((x[1:] - x[npos[1:, 0]]) ** 2).sum(1)
    assert (np.diff(d_near) <= 0.0).all()
    # brute force on a sample of positions
    rng = np.random.default_rng(0)
    for k in np.concatenate([[1, 2, m, m + 1], rng.integers(m + 2, n, 150)]):
        d2 = ((x[:k] - x[k]) ** 2).sum(1)
        thr = np.partition(d2, min(m, k) - 1)[min(m, k) - 1]
        cand = np.flatnonzero(d2 <= thr)
        want = cand[np.lexsort((cand, d2[cand]))][:m]
        assert np.array_equal(npos[k, :cnt[k]], want), k
    vals = workloads.vecchia_field(c, o, nb, cnt, m)
    models = workloads.vecchia_models(8)
    a, sa = p.loglik_batch(vals, models)
    b, sb = p.loglik_batch(vals, models)
    assert (sa == 0).all() and np.array_equal(a, b) and np.array_equal(sa, sb)


def test_full_size_loglik_against_oracle(oracle):
    import concurrent.futures as cf
    c = workloads.vecchia_c3()
    m = 30
    p = gh.vecchia_plan(c, m)
    o, nb, cnt = p.arrays()
    vals = workloads.vecchia_field(c, o, nb, cnt, m)
    models = workloads.vecchia_models(64)
    picks = [0, 37]          # two of the 64 C3 parameter vectors
    print(f"vecchia full size vs oracle: n={len(c)} m={m} models {picks}", flush=True)
    ll, st = p.loglik_batch(vals, models)
    assert (st == 0).all()
    with cf.ThreadPoolExecutor(max_workers=len(picks)) as ex:
        refs = list(ex.map(lambda i: oracle.vecchia_loglik_plan(c, vals, models[i], o, nb, cnt),
                           picks))
    for i, (rc, oll) in zip(picks, refs):
        assert rc == 0
        assert abs(ll[i] - oll) <= 1e-10 * abs(oll), (i, ll[i], oll)
