"""Seeded sweep over random Vecchia plans and models against the oracle.

Each case draws a dimension, a point set (uniform, clustered, or snapped to
a lattice so that distances tie), a conditioning size m from 1 to 64, the
maxmin or a given order, and a batch of models of both the lean and the
general class; a second sweep checks the likelihood gradient of random
lean models.  Bars as in the Vecchia tests: the plan bit-exact, every
log-likelihood within 1e-10 relative, status codes equal, gradients within
1e-8 normwise.
"""
import numpy as np
import pytest

import geostat_hip as gh
from test_gpu_vecchia import plan_equal

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
KINDS = [gh.ModelKind.SPHERICAL, gh.ModelKind.EXPONENTIAL, gh.ModelKind.GAUSSIAN,
         gh.ModelKind.MATERN15, gh.ModelKind.MATERN25, gh.ModelKind.Matern(1.2),
         gh.ModelKind.Stable(1.2), gh.ModelKind.WAVE]


def _case(seed):
    r = np.random.default_rng(5000 + seed)
    D = 2 if seed % 3 else 3
    n = int(r.integers(100, 2500))
    shape = seed % 4
    if shape == 0:                      # uniform
        c = r.uniform(0.0, 1.0, size=(n, D))
    elif shape == 1:                    # clusters
        centres = r.uniform(0.0, 1.0, size=(8, D))
        c = centres[r.integers(8, size=n)] + 0.03 * r.standard_normal((n, D))
    else:                               # lattice-snapped: tied distances
        c = np.round(r.uniform(0.0, 1.0, size=(n, D)) * 40.0) / 40.0
    c = np.unique(c, axis=0)
    r.shuffle(c)
    n = len(c)
    m = int(r.integers(1, 65))
    order = r.permutation(n) if r.uniform() < 0.3 else None
    models = []
    for _ in range(int(r.integers(1, 9))):
        kind = KINDS[int(r.integers(len(KINDS)))]
        models.append(gh.VariogramModel(float(r.uniform(0.01, 0.1)),
                                        [gh.Structure(kind, float(r.uniform(0.5, 2.0)),
                                                      float(r.uniform(0.03, 0.4)))]))
    v = np.sin(c[:, 0] * 7.0) + np.cos(c[:, 1] * 5.0) + 0.1 * r.standard_normal(n)
    return D, c, v, m, order, models


@pytest.mark.parametrize("seed", range(32))
def test_random_plan_and_models_match_oracle(oracle, seed):
    D, c, v, m, order, models = _case(seed)
    print(f"vecchia sweep seed {seed}: D={D} n={len(c)} m={m} given order={order is not None} "
          f"models={len(models)}", flush=True)
    p, o, nb, cnt = plan_equal(oracle, c, m, order)
    ll, st = p.loglik_batch(v, models)
    for i, mdl in enumerate(models):
        rc, oll = oracle.vecchia_loglik_plan(c, v, mdl, o, nb, cnt)
        if rc != 0:
            assert st[i] != 0, (i, rc, st[i])
            continue
        assert st[i] == 0, (i, st[i])
        assert abs(ll[i] - oll) <= TOL64 * abs(oll), (i, ll[i], oll)


@pytest.mark.parametrize("seed", range(16))
def test_random_plan_gradient_matches_oracle(oracle, seed):
    from test_gpu_vecchia import GRAD_KINDS, _check_grad
    D, c, v, _, order, _ = _case(100 + seed)
    r = np.random.default_rng(7000 + seed)
    m = int(r.integers(1, 33))          # the gradient kernels take sets up to 32
    models = [gh.VariogramModel(float(r.uniform(0.01, 0.3)),
                                [gh.Structure(GRAD_KINDS[int(r.integers(len(GRAD_KINDS)))],
                                              float(r.uniform(0.3, 2.0)),
                                              float(r.uniform(0.03, 0.4)))])
              for _ in range(int(r.integers(1, 9)))]
    print(f"gradient sweep seed {seed}: D={D} n={len(c)} m={m} models={len(models)}", flush=True)
    p, o, nb, cnt = plan_equal(oracle, c, m, order)
    _check_grad(oracle, c, v, o, nb, cnt, p, models)
