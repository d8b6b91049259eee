"""Seeded sweep over random predictor configurations against the oracle.

Each case draws a dimension, a data set, a nested model (kinds of both the
lean and the general class, isotropic or with random anisotropy), a method
(SK / OK / UK degree 1 or 2) and a neighbourhood (k nearest, radius with a
minimum, octant quotas, anisotropic search); targets reach past the data,
so short neighbourhoods fail.  Bars as in the kriging tests:
neighbour lists bit-exact, estimates within 1e-10 (or closer than the
oracle to the 40-digit solution), per-target status codes equal.
"""
import numpy as np
import pytest

import geostat_hip as gh
from test_gpu_kriging import check_neighbors, check_predict

pytestmark = pytest.mark.gpu

LEAN = [gh.ModelKind.SPHERICAL, gh.ModelKind.EXPONENTIAL, gh.ModelKind.GAUSSIAN,
        gh.ModelKind.MATERN15, gh.ModelKind.MATERN25]
GENERAL = [gh.ModelKind.Matern(0.8), gh.ModelKind.Stable(1.5), gh.ModelKind.HOLE,
           gh.ModelKind.WAVE]


def _anis(r, D):
    if D == 2:
        return gh.Anisotropy(float(r.uniform(0, 180)), float(r.uniform(0.3, 1.0)))
    return gh.Anisotropy(float(r.uniform(0, 180)), float(r.uniform(0.3, 1.0)),
                         ratio_z=float(r.uniform(0.2, 1.0)), dip_deg=float(r.uniform(-30, 30)),
                         rake_deg=float(r.uniform(-20, 20)))


def _case(seed):
    r = np.random.default_rng(1000 + seed)
    D = 2 if seed % 3 else 3
    n = int(r.integers(300, 4000))
    ext = np.array([1000.0, 1000.0, 200.0][:D])
    c = r.uniform(0.0, 1.0, size=(n, D)) * ext
    v = np.sin(c[:, 0] / 150.0) + np.cos(c[:, 1] / 200.0) + 0.2 * r.standard_normal(n)
    if D == 3:
        v += c[:, 2] / 300.0
    kinds = LEAN if seed % 2 == 0 else LEAN + GENERAL
    structs = []
    shared = _anis(r, D) if r.uniform() < 0.5 else None
    for _ in range(int(r.integers(1, 3))):
        kind = kinds[int(r.integers(len(kinds)))]
        an = shared if shared is not None else (_anis(r, D) if r.uniform() < 0.4 else None)
        structs.append(gh.Structure(kind, float(r.uniform(0.3, 1.0)), float(r.uniform(80, 500)), an))
    model = gh.VariogramModel(float(r.uniform(0.02, 0.2)), structs)
    method = [gh.KrigingMethod.Simple(0.3), gh.KrigingMethod.Ordinary(),
              gh.KrigingMethod.Universal(1), gh.KrigingMethod.Universal(2)][int(r.integers(4))]
    m_drift = {0: 0, 1: 1}.get(method.kind, 1 + D if method.degree == 1 else 1 + D + D * (D + 1) // 2)
    mode = int(r.integers(3))
    if mode == 0:       # k nearest
        cfg = gh.KrigingConfig(method=method, max_neighbors=int(r.integers(max(4, 3 * m_drift), 65)))
    elif mode == 1:     # radius-limited with a minimum
        cfg = gh.KrigingConfig(method=method, max_neighbors=int(r.integers(max(8, 3 * m_drift), 49)),
                               search_radius=float(r.uniform(30, 250)),
                               min_neighbors=int(r.integers(max(1, m_drift), max(2, m_drift) + 4)))
    else:               # octant quotas, possibly anisotropic search
        q = int(r.integers(max(1, (3 * m_drift + (1 << D) - 1) >> D), (64 >> D) + 1))
        cfg = gh.KrigingConfig(method=method, max_per_octant=q,
                               search_radius=float(r.uniform(60, 400)),
                               min_neighbors=max(1, m_drift),
                               anisotropic_search=_anis(r, D) if r.uniform() < 0.5 else None)
    t = r.uniform(-0.2, 1.2, size=(300, D)) * ext
    t = np.concatenate([t, c[r.choice(n, 20, replace=False)]])
    return D, c, v, model, cfg, t


@pytest.mark.parametrize("seed", range(256))
def test_random_configuration_matches_oracle(oracle, seed):
    D, c, v, model, cfg, t = _case(seed)
    print(f"sweep seed {seed}: D={D} n={len(c)} structures={len(model.structures)} "
          f"method={cfg.method.kind}/{cfg.method.degree} k={cfg.max_neighbors} "
          f"octant={cfg.max_per_octant} radius={cfg.search_radius}", flush=True)
    check_neighbors(oracle, c, v, model, cfg, t)
    check_predict(oracle, c, v, model, cfg, t)
