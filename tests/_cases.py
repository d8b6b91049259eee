"""Shared fixtures restating the reference's own test inputs.

Each builder names the reference test it comes from (paths under
crates/geostat-core/src/ unless stated).  They return plain numpy data plus
product-API model/config objects, which the oracle wrapper also accepts.
"""
from __future__ import annotations

import numpy as np

from geostat_hip import (Anisotropy, KrigingConfig, KrigingMethod, ModelKind, Structure,
                         VariogramModel)


def sample_data():
    """kriging.rs:845-851"""
    c = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [1.0, 1.0], [0.4, 0.6]])
    v = np.array([1.0, 2.0, 1.5, 2.5, 1.7])
    return c, v


def sample_model():
    """kriging.rs:853-855: 0.05 nugget + spherical(0.95, 2.0)"""
    return VariogramModel(0.05, [Structure(ModelKind.SPHERICAL, 0.95, 2.0)])


def cfg(**kw):
    return KrigingConfig(**kw)


def rng_points(oracle, n, seed, scale=100.0):
    """search.rs:355-361 random_points: [U*100, U*100]."""
    r = oracle.Rng(seed)
    return np.array([[r.uniform() * scale, r.uniform() * scale] for _ in range(n)])


def vecchia_field(oracle, n, seed):
    """vecchia.rs:1943-1954 field(): x, y = U*100; v = sin(x/25) + cos(y/30) + 0.1 N."""
    r = oracle.Rng(seed)
    c, v = [], []
    for _ in range(n):
        x = r.uniform() * 100.0
        y = r.uniform() * 100.0
        c.append([x, y])
        v.append(np.sin(x / 25.0) + np.cos(y / 30.0) + 0.1 * r.normal())
    return np.array(c), np.array(v)


def octant_cluster():
    """kriging.rs:960-1011 octant_search_balances_clustered_neighbors"""
    c = [[1.0, 0.1], [1.1, -0.1], [1.2, 0.2], [1.3, -0.2], [1.4, 0.05]]
    v = [10.0] * 5
    c += [[-2.0, 0.5], [-2.0, -0.5], [0.5, -2.0], [0.5, 2.0]]
    v += [0.0] * 4
    return np.array(c), np.array(v)


def octant_cluster_rotated():
    """kriging.rs:1103-1157"""
    c = [[0.1, 1.0], [-0.1, 1.1], [0.2, 1.2], [-0.2, 1.3], [0.05, 1.4]]
    v = [10.0] * 5
    c += [[0.5, -2.0], [-0.5, -2.0], [-2.0, 0.5], [2.0, 0.5]]
    v += [0.0] * 4
    return np.array(c), np.array(v)


def lattice(nx, ny, step=1.0, jitter_seed=None):
    """Integer lattice: k-th neighbour distances tie, exercising the kd-tree
    traversal-order rule of search.rs:15-22, 86-98."""
    xs, ys = np.meshgrid(np.arange(nx) * step, np.arange(ny) * step)
    c = np.stack([xs.ravel(), ys.ravel()], axis=1)
    v = np.sin(c[:, 0] / 3.0) + np.cos(c[:, 1] / 4.0)
    return c, v


def drift_basis(coords, method, p):
    """kriging.rs:340-347, 410-439: [1, xn, (xn_a xn_b, a <= b)] with
    xn = (p - centre) / scale over the data bounding box."""
    if method.kind == 0:
        return []
    if method.kind == 1:
        return [1.0]
    mn, mx = coords.min(0), coords.max(0)
    centre = 0.5 * (mn + mx)
    scale = max(0.5 * float((mx - mn).max()), 2.2250738585072014e-308)
    xn = (np.asarray(p) - centre) / scale
    out = [1.0] + list(xn)
    if method.degree == 2:
        D = len(xn)
        out += [xn[a] * xn[b] for a in range(D) for b in range(a, D)]
    return out


def exact_estimate(oracle, coords, values, model, method, target, nb, digits=40):
    """The bordered kriging system (kriging.rs:527-644) with entries from the
    oracle's gamma, solved in `digits`-digit arithmetic: the exact answer the
    double-precision solvers approximate.  Returns (value, variance,
    multiplier), the multiplier being the solution's first drift entry (NaN
    without drift)."""
    import mpmath as mp
    mp.mp.dps = digits
    c0 = model.total_sill()
    n = len(nb)
    F = [drift_basis(coords, method, coords[i]) for i in nb]
    m = len(F[0]) if F else 0
    A = mp.zeros(n + m, n + m)
    b = mp.zeros(n + m, 1)
    for a in range(n):
        for q in range(n):
            A[a, q] = c0 if a == q else c0 - oracle.gamma_dh(model, coords[nb[a]] - coords[nb[q]])
        b[a] = c0 - oracle.gamma_dh(model, coords[nb[a]] - target)
        for k in range(m):
            A[a, n + k] = F[a][k]
            A[n + k, a] = F[a][k]
    for k, fk in enumerate(drift_basis(coords, method, target)):
        b[n + k] = fk
    w = mp.lu_solve(A, b)
    if method.kind == 0:
        val = method.mean + sum(w[a] * (values[nb[a]] - method.mean) for a in range(n))
        var = c0 - sum(w[a] * b[a] for a in range(n))
    else:
        val = sum(w[a] * values[nb[a]] for a in range(n))
        var = c0 - sum(w[a] * b[a] for a in range(n + m))
    lag = float(w[n]) if m > 0 else float("nan")
    return float(val), max(float(var), 0.0), lag


MODELS_2D = {
    "sph": VariogramModel(0.1, [Structure(ModelKind.SPHERICAL, 1.0, 250.0)]),
    "exp": VariogramModel(0.05, [Structure(ModelKind.EXPONENTIAL, 1.0, 150.0)]),
    "gau": VariogramModel(0.05, [Structure(ModelKind.GAUSSIAN, 1.0, 200.0)]),
    "mat15_aniso": VariogramModel(0.05, [Structure(ModelKind.MATERN15, 1.0, 300.0,
                                                   Anisotropy(30.0, 0.5))]),
    "mat25": VariogramModel(0.02, [Structure(ModelKind.MATERN25, 1.0, 200.0)]),
    "matern_nu": VariogramModel(0.05, [Structure(ModelKind.Matern(0.8), 1.0, 200.0)]),
    "circular": VariogramModel(0.05, [Structure(ModelKind.CIRCULAR, 1.0, 250.0)]),
    "stable": VariogramModel(0.05, [Structure(ModelKind.Stable(1.5), 1.0, 200.0)]),
    "hole": VariogramModel(0.1, [Structure(ModelKind.HOLE, 1.0, 80.0)]),
    "wave": VariogramModel(0.1, [Structure(ModelKind.WAVE, 1.0, 200.0)]),
    "nested": VariogramModel(0.05, [Structure(ModelKind.EXPONENTIAL, 0.5, 100.0),
                                    Structure(ModelKind.SPHERICAL, 0.5, 400.0,
                                              Anisotropy(120.0, 0.4))]),
}
