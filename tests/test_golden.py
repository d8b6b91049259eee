"""Golden fixtures (tests/golden, written by make_golden.py from the oracle on
the reference's workloads).  CPU: the oracle and the synthetic-data
generator still reproduce them.  GPU: neighbour lists and plans bit-exact,
FP64 estimates and log-likelihoods within 1e-10 relative."""
import math
from pathlib import Path

import numpy as np
import pytest

import _cases as cs
import geostat_hip as gh
from geostat_hip import workloads

G = Path(__file__).resolve().parent / "golden"


def load(name):
    return np.load(G / name, allow_pickle=False)


def close(a, b, floor, tol=1e-10):
    a, b = np.asarray(a), np.asarray(b)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    m = ~np.isnan(b)
    if m.any():
        assert np.abs(a[m] - b[m]).max() <= tol * max(np.abs(b[m]).max(), floor)


# ---- CPU: the fixtures are reproducible ------------------------------------
def test_rng_and_generator_fixture(oracle):
    f = load("rng.npz")
    for seed in (0, 42, 2**63 + 5):
        g = oracle.Rng(seed)
        assert np.array_equal(np.array([g.next_u64() for _ in range(16)], dtype=np.uint64),
                              f[f"u64_{seed}"])
        g = oracle.Rng(seed)
        assert np.array_equal(np.array([g.uniform() for _ in range(8)]), f[f"uniform_{seed}"])
        g = oracle.Rng(seed)
        assert np.abs(np.array([g.normal() for _ in range(9)]) - f[f"normal_{seed}"]).max() < 1e-15
    c, v = workloads.points(42, 64, [1000.0, 1000.0])
    assert np.abs(np.concatenate([c.ravel(), v]) - f["points42_head"]).max() < 1e-12


def _krige_inputs(name):
    if name == "c0.npz":
        w = workloads.ok2d_small()
        return w.coords, w.values, w.model, w.config, 20
    if name == "c1_slice.npz":
        w = workloads.ok2d_large()
        return w.coords, w.values, w.model, w.config, 32
    w = workloads.uk3d_octant()
    return w.coords, w.values, w.model, w.config, 64


@pytest.mark.parametrize("name", ["c0.npz", "c1_slice.npz", "c2_slice.npz"])
def test_oracle_reproduces_kriging_fixtures(oracle, name):
    f = load(name)
    c, v, m, cfg, kmax = _krige_inputs(name)
    t = f["targets"]
    rc, idx, cnt = oracle.neighbors(c, v, m, cfg, t, kmax)
    assert rc == 0 and np.array_equal(idx, f["idx"]) and np.array_equal(cnt, f["cnt"])
    rc, val, var, lag, st, _ = oracle.predict_many(c, v, m, cfg, t)
    assert np.array_equal(st, f["status"])
    c0 = m.total_sill()
    close(val, f["value"], 1e-3 * math.sqrt(c0), 1e-13)
    close(var, f["variance"], 1e-3 * c0, 1e-13)


def test_oracle_reproduces_tie_and_vecchia_fixtures(oracle):
    f = load("ties.npz")
    c, v = cs.lattice(23, 19)
    m = gh.VariogramModel(0.1, [gh.Structure(gh.ModelKind.SPHERICAL, 1.0, 6.0)])
    for k in (4, 8, 13):
        t = f[f"k{k}_targets"]
        rc, idx, cnt = oracle.neighbors(c, v, m, gh.KrigingConfig(max_neighbors=k), t, k)
        assert np.array_equal(idx, f[f"k{k}_idx"]) and np.array_equal(cnt, f[f"k{k}_cnt"])
    f = load("vecchia.npz")
    c, v = cs.vecchia_field(oracle, 2000, 7)
    rc, order, nb, cnt = oracle.vecchia_plan(c, 30)
    assert np.array_equal(order, f["order"]) and np.array_equal(nb, f["nb"])
    models = workloads.vecchia_models(4, seed=99)
    lls = [oracle.vecchia_loglik_plan(c, v, mm, order, nb, cnt)[1] for mm in models]
    close(lls, f["loglik"], 1.0, 1e-13)


# ---- GPU against the fixtures ----------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c0.npz", "c1_slice.npz", "c2_slice.npz"])
def test_gpu_matches_kriging_fixtures(name):
    f = load(name)
    c, v, m, cfg, kmax = _krige_inputs(name)
    k = gh.Kriging(c, v, m, cfg)
    t = f["targets"]
    idx, cnt = k.neighbors(t)
    w = f["idx"].shape[1]
    assert np.array_equal(cnt, f["cnt"])
    assert np.array_equal(idx[:, :w], f["idx"])
    val, var, lag, st = k.predict_many(t)
    assert np.array_equal(st, f["status"])
    c0 = m.total_sill()
    close(val, f["value"], 1e-3 * math.sqrt(c0))
    close(var, f["variance"], 1e-3 * c0)


@pytest.mark.gpu
def test_gpu_matches_tie_and_vecchia_fixtures(oracle):
    f = load("ties.npz")
    c, v = cs.lattice(23, 19)
    m = gh.VariogramModel(0.1, [gh.Structure(gh.ModelKind.SPHERICAL, 1.0, 6.0)])
    for kk in (4, 8, 13):
        k = gh.Kriging(c, v, m, gh.KrigingConfig(max_neighbors=kk))
        idx, cnt = k.neighbors(f[f"k{kk}_targets"])
        assert np.array_equal(idx, f[f"k{kk}_idx"]) and np.array_equal(cnt, f[f"k{kk}_cnt"])
        val, var, _, st = k.predict_many(f[f"k{kk}_targets"])
        close(val, f[f"k{kk}_value"], 1e-3)
        close(var, f[f"k{kk}_variance"], 1e-3)
    f = load("vecchia.npz")
    c, v = cs.vecchia_field(oracle, 2000, 7)
    p = gh.vecchia_plan(c, 30)
    order, nb, cnt = p.arrays()
    assert np.array_equal(order, f["order"]) and np.array_equal(nb, f["nb"])
    ll, st = p.loglik_batch(v, workloads.vecchia_models(4, seed=99))
    assert (st == 0).all()
    close(ll, f["loglik"], 1.0)
