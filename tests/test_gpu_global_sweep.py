"""Seeded sweep over random global-neighbourhood predictors (n from 65 to
600: one shared LU on the GPU) against the oracle, with plain, measurement
error, external drift and block variants, random nested models and 2-D / 3-D
data.  Bars as in the global tests: statuses equal, value, variance and
multiplier within 1e-10 relative."""
import numpy as np
import pytest

import geostat_hip as gh
from test_gpu_global import compare

pytestmark = pytest.mark.gpu

KINDS = [gh.ModelKind.SPHERICAL, gh.ModelKind.EXPONENTIAL, gh.ModelKind.GAUSSIAN,
         gh.ModelKind.MATERN15, gh.ModelKind.MATERN25, gh.ModelKind.Matern(0.8),
         gh.ModelKind.Stable(1.5), gh.ModelKind.HOLE]


@pytest.mark.parametrize("seed", range(16))
def test_random_global_predictor_matches_oracle(oracle, seed):
    r = np.random.default_rng(11000 + seed)
    D = 2 if seed % 3 else 3
    n = int(r.integers(65, 600))
    ext = np.array([1000.0, 1000.0, 150.0][:D])
    c = r.uniform(0.0, 1.0, size=(n, D)) * ext
    v = np.sin(c[:, 0] / 150.0) + np.cos(c[:, 1] / 120.0) + 0.2 * r.standard_normal(n)
    structs = []
    for _ in range(int(r.integers(1, 3))):
        kind = KINDS[int(r.integers(len(KINDS)))]
        an = None
        if r.uniform() < 0.4:
            an = (gh.Anisotropy(float(r.uniform(0, 180)), float(r.uniform(0.4, 1.0))) if D == 2 else
                  gh.Anisotropy(float(r.uniform(0, 180)), float(r.uniform(0.4, 1.0)),
                                ratio_z=float(r.uniform(0.3, 1.0)), dip_deg=float(r.uniform(-20, 20))))
        structs.append(gh.Structure(kind, float(r.uniform(0.4, 1.0)), float(r.uniform(100, 500)), an))
    model = gh.VariogramModel(float(r.uniform(0.05, 0.2)), structs)
    t = r.uniform(-0.1, 1.1, size=(100, D)) * ext
    t[:10] = c[:10]
    variant = ("plain", "error", "drift", "block")[seed % 4]
    print(f"global sweep seed {seed}: {variant} D={D} n={n} structures={len(structs)}", flush=True)
    if variant == "drift":
        ang = [np.deg2rad(70.0 * j + 10.0) for j in range(2)]
        cov = lambda p: np.column_stack([np.sin((p[:, 0] * np.cos(a) + p[:, 1] * np.sin(a)) / 200.0)
                                         for a in ang])
        dd, td = cov(c), cov(t)
        v = v + dd.sum(1)
        cfg = gh.KrigingConfig(method=gh.KrigingMethod.ExternalDrift(2))
        k = gh.Kriging.with_external_drift(c, v, model, cfg, dd)
        compare(model, k.predict_many_with_drift(t, td),
                oracle.predict_many_ex(c, v, model, cfg, t, drift=dd, tdrift=td))
        return
    method = [gh.KrigingMethod.Simple(0.2), gh.KrigingMethod.Ordinary(),
              gh.KrigingMethod.Universal(1), gh.KrigingMethod.Universal(2)][int(r.integers(4))]
    if variant == "block":
        method = [gh.KrigingMethod.Simple(0.2), gh.KrigingMethod.Ordinary()][int(r.integers(2))]
    cfg = gh.KrigingConfig(method=method)
    if variant == "plain":
        k = gh.Kriging(c, v, model, cfg)
        compare(model, k.predict_many(t), oracle.predict_many(c, v, model, cfg, t))
    elif variant == "error":
        err = np.abs(r.standard_normal(n)) * 0.05
        k = gh.Kriging.with_measurement_error(c, v, model, cfg, err)
        compare(model, k.predict_many(t), oracle.predict_many_ex(c, v, model, cfg, t, errors=err))
    else:
        offs = r.uniform(-15.0, 15.0, size=(int(r.integers(1, 9)), D))
        k = gh.Kriging(c, v, model, cfg)
        compare(model, k.predict_block_many(t, offs),
                oracle.predict_many_ex(c, v, model, cfg, t, offsets=offs, block=True))
