"""Properties of the 3-D octant predictor at its full size (10^6 data
points, 200 x 200 x 100 = 4,000,000 targets), where the oracle is run on a
sample only: every target succeeds, results are deterministic and
independent of the batch they come in, a 500-target sample matches the
oracle (lists bit-exact, estimates within 1e-10), and data locations are
reproduced."""
import numpy as np
import pytest

import geostat_hip as gh
from geostat_hip import workloads
from test_gpu_kriging import check_neighbors, check_predict

pytestmark = pytest.mark.gpu


def test_uk3d_octant_full_size_properties(oracle):
    w = workloads.uk3d_octant()
    k = gh.Kriging(w.coords, w.values, w.model, w.config)
    t = w.targets()
    print(f"uk3d full size: n={len(w.coords)} T={len(t)}", flush=True)
    v1, var1, lag1, st1 = k.predict_many(t)
    assert (st1 == 0).all() and np.isfinite(v1).all() and (var1 >= 0).all()
    v2, var2, _, _ = k.predict_many(t)
    assert np.array_equal(v1, v2) and np.array_equal(var1, var2)
    sel = np.random.default_rng(1).choice(len(t), 5000, replace=False)
    v3, var3, _, _ = k.predict_many(t[sel])
    assert np.array_equal(v3, v1[sel]) and np.array_equal(var3, var1[sel])
    sample = t[np.random.default_rng(2).choice(len(t), 500, replace=False)]
    check_neighbors(oracle, w.coords, w.values, w.model, w.config, sample)
    check_predict(oracle, w.coords, w.values, w.model, w.config, sample)
    # nugget inside c0: a data location returns its datum with zero variance
    v4, var4, _, st4 = k.predict_many(w.coords[:1000])
    assert (st4 == 0).all()
    assert np.abs(v4 - w.values[:1000]).max() < 1e-8 and var4.max() < 1e-8
