"""Model JSON (serde form) and the reference's Python entry points.

CPU: JSON parsing / writing against the model files the reference's own
scripts write (validation/v02_gstat.R:75-80, cli tests), round trips and
validation.  GPU: krige / krige_grid / krige_3d / loo_cv / vecchia_loglik
against the oracle.
"""
import json
import math

import numpy as np
import pytest

import geostat_hip as gh
from geostat_hip import frontend as fe

ANISO_JSON = ('{"nugget": 0.05, "structures": [{"kind": "spherical", "sill": 0.55, '
              '"range": 900.0, "anis": {"azimuth_deg": 30.0, "ratio": 0.5}}]}')


def test_model_json_reference_files():
    m = gh.VariogramModel.from_json(ANISO_JSON)
    assert m.nugget == 0.05 and len(m.structures) == 1
    s = m.structures[0]
    assert s.kind == gh.ModelKind.SPHERICAL and s.sill == 0.55 and s.range == 900.0
    assert s.anis == gh.Anisotropy(30.0, 0.5, 1.0, 0.0, 0.0)
    # cli.rs:363
    m2 = gh.VariogramModel.from_json(
        '{"nugget":0.1,"structures":[{"kind":"spherical","sill":0.9,"range":10.0}]}')
    assert m2.to_json() == '{"nugget":0.1,"structures":[{"kind":"spherical","sill":0.9,"range":10.0}]}'


def test_model_json_parametrised_kinds_and_round_trip():
    m = gh.VariogramModel(0.0, [
        gh.Structure(gh.ModelKind.Matern(1.2), 0.5, 100.0),
        gh.Structure(gh.ModelKind.Stable(1.5), 0.25, 1e-7,
                     gh.Anisotropy(12.5, 0.3, 0.2, 15.0, -5.0)),
        gh.Structure(gh.ModelKind.MATERN25, 1.0, 3.0)])
    txt = m.to_json()
    d = json.loads(txt)
    assert d["structures"][0]["kind"] == {"matern": 1.2}
    assert d["structures"][1]["kind"] == {"stable": 1.5}
    assert d["structures"][2]["kind"] == "matern25"
    assert "1e-7" in txt and "anis" not in json.dumps(d["structures"][0])
    assert d["structures"][1]["anis"] == {"azimuth_deg": 12.5, "ratio": 0.3, "ratio_z": 0.2,
                                          "dip_deg": 15.0, "rake_deg": -5.0}
    assert gh.VariogramModel.from_json(txt) == m


@pytest.mark.parametrize("bad", [
    '{"nugget": -0.1, "structures": [{"kind": "spherical", "sill": 1.0, "range": 1.0}]}',
    '{"nugget": 0.0, "structures": [{"kind": "spherical", "sill": 1.0, "range": 0.0}]}',
    '{"nugget": 0.0, "structures": [{"kind": {"matern": 20.0}, "sill": 1.0, "range": 1.0}]}',
    '{"nugget": 0.0, "structures": [{"kind": "cubic", "sill": 1.0, "range": 1.0}]}',
    '{"nugget": 0.0, "structures": []}',
    '{"nugget": 0.0}',
])
def test_model_json_rejections(bad):
    with pytest.raises(ValueError):
        gh.VariogramModel.from_json(bad)


def test_read_write_model(tmp_path):
    p = tmp_path / "model.json"
    p.write_text(ANISO_JSON)
    m = fe.read_model(p)
    fe.write_model(tmp_path / "out.json", m)
    assert fe.read_model(tmp_path / "out.json") == m


@pytest.mark.gpu
def test_krige_and_grid_match_oracle(oracle):
    from geostat_hip import workloads
    c, v = workloads.points(81, 2000, [1000.0, 1000.0])
    m = gh.VariogramModel.from_json(ANISO_JSON)
    t = workloads.points(82, 500, [1000.0, 1000.0])[0]
    val, var = fe.krige(c[:, 0], c[:, 1], v, m, t[:, 0], t[:, 1], max_neighbors=16,
                        search_azimuth=30.0, search_ratio=0.5)
    cfg = gh.KrigingConfig(max_neighbors=16, anisotropic_search=gh.Anisotropy(30.0, 0.5))
    rc, ov, ovar, _, _, _ = oracle.predict_many(c, v, m, cfg, t)
    assert np.abs(val - ov).max() < 1e-10 * np.abs(ov).max()
    assert np.abs(var - ovar).max() < 1e-10 * np.abs(ovar).max()
    # simple kriging defaults its mean to the data mean
    val, _ = fe.krige(c[:, 0], c[:, 1], v, m, t[:, 0], t[:, 1], method="simple",
                      max_neighbors=16)
    mean = -0.0
    for x in v.tolist():
        mean += x
    mean /= len(v)
    ov = oracle.predict_many(c, v, m, gh.KrigingConfig(method=gh.KrigingMethod.Simple(mean),
                                                       max_neighbors=16), t)[1]
    assert np.abs(val - ov).max() < 1e-10 * np.abs(ov).max()
    gv, gvar = fe.krige_grid(c[:, 0], c[:, 1], v, m, (0.0, 0.0, 1000.0, 1000.0), 20, 10,
                             max_neighbors=16)
    grid = gh.Grid2D.from_bbox([0.0, 0.0], [1000.0, 1000.0], 20, 10)
    ov = oracle.predict_many(c, v, m, gh.KrigingConfig(max_neighbors=16), grid.centers())[1]
    assert np.abs(gv - ov).max() < 1e-10 * np.abs(ov).max()
    bv, _ = fe.krige_grid(c[:, 0], c[:, 1], v, m, (0.0, 0.0, 1000.0, 1000.0), 20, 10,
                          max_neighbors=16, block=(50.0, 100.0), block_discr=(2, 3))
    ob = oracle.predict_many_ex(c, v, m, gh.KrigingConfig(max_neighbors=16), grid.centers(),
                                offsets=gh.block_offsets([50.0, 100.0], [2, 3]), block=True)[1]
    assert np.abs(bv - ob).max() < 1e-10 * np.abs(ob).max()
    with pytest.raises(ValueError):
        fe.krige(c[:, 0], c[:, 1], v, m, t[:, 0], t[:5, 1])
    with pytest.raises(ValueError):
        fe.krige(c[:, 0], c[:, 1], v, m, t[:, 0], t[:, 1], method="kriging")


@pytest.mark.gpu
def test_krige_3d_loo_cv_and_vecchia_loglik(oracle):
    from geostat_hip import workloads
    c, v = workloads.points(83, 1500, [500.0, 500.0, 50.0], recipe=1)
    m = gh.VariogramModel(0.1, [gh.Structure(gh.ModelKind.EXPONENTIAL, 1.0, 150.0)])
    t = workloads.points(84, 300, [500.0, 500.0, 50.0])[0]
    val, var = fe.krige_3d(c[:, 0], c[:, 1], c[:, 2], v, m, t[:, 0], t[:, 1], t[:, 2],
                           method="universal", max_neighbors=20)
    cfg = gh.KrigingConfig(method=gh.KrigingMethod.Universal(1), max_neighbors=20)
    ov = oracle.predict_many(c, v, m, cfg, t)[1]
    assert np.abs(val - ov).max() < 1e-10 * np.abs(ov).max()
    r = oracle.Rng(11)
    x, y, z = [], [], []
    for _ in range(120):
        a = r.uniform() * 100.0
        b = r.uniform() * 100.0
        x.append(a)
        y.append(b)
        z.append(math.sin(a / 20.0) + math.cos(b / 20.0))
    mm = gh.VariogramModel(0.01, [gh.Structure(gh.ModelKind.SPHERICAL, 1.0, 50.0)])
    d = fe.loo_cv(x, y, z, mm)
    assert d["rmse"] < 0.5 * np.std(z) and np.isfinite(d["msdr"]) and len(d["predicted"]) == 120
    d5 = fe.loo_cv(x, y, z, mm, folds=5, seed=42)
    fo, mem = oracle.fold_assignment(120, 5, 42)
    rc, p, _ = oracle.cv_with_folds(np.column_stack([x, y]), np.array(z), mm, gh.KrigingConfig(),
                                    fo, mem)
    assert rc == 0 and np.abs(d5["predicted"] - p).max() < 1e-10 * np.abs(p).max()
    with pytest.raises(ValueError):
        fe.loo_cv(x, y, z, mm, folds=5, blocks=(2, 2))
    c2, v2 = c[:, :2], v
    ll = fe.vecchia_loglik(c2[:, 0], c2[:, 1], v2, m, m=10)
    rc, o, nb, cnt = oracle.vecchia_plan(c2, 10)
    rc2, oll = oracle.vecchia_loglik_plan(c2, v2, m, o, nb, cnt)
    assert abs(ll - oll) <= 1e-10 * abs(oll)
