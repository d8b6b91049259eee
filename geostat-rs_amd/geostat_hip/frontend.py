"""The reference's Python entry points for this path (python/src/lib.rs),
over the GPU predictor: same names, argument meaning and defaults, and the
same (predictions, variances) returns; failed targets are NaN.  Errors are
raised as ValueError subclasses (GeostatError), as the reference raises
ValueError.  Model files use the serde JSON form (io_utils.rs:274-289)."""
from __future__ import annotations

from typing import Optional, Sequence, Tuple

import numpy as np

from . import (Anisotropy, Grid2D, InvalidParameter, Kriging, KrigingConfig, KrigingMethod,
               VariogramModel, vecchia_loglik as _vecchia_loglik)
from . import validation as _cv

ArrayPair = Tuple[np.ndarray, np.ndarray]


def read_model(path) -> VariogramModel:
    """read_model (io_utils.rs:274-281): JSON file, re-validated."""
    with open(path, "r", encoding="utf-8") as f:
        return VariogramModel.from_json(f.read())


def write_model(path, model: VariogramModel) -> None:
    """write_model (io_utils.rs:284-289)."""
    with open(path, "w", encoding="utf-8") as f:
        f.write(model.to_json())


def _coords2(x, y, values):
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    v = np.asarray(values, dtype=np.float64)
    if not (len(x) == len(y) == len(v)):   # PointSet::from_xyz (data.rs:36-55)
        from . import DimensionMismatch
        raise DimensionMismatch(f"x, y and values differ in length ({len(x)}, {len(y)}, {len(v)})")
    return np.column_stack([x, y]), v


def _method(method: str, mean, degree: int, values) -> KrigingMethod:
    """build_method (python/src/lib.rs:104-120)."""
    if method == "ordinary":
        return KrigingMethod.Ordinary()
    if method == "simple":
        if mean is None:
            s = -0.0
            for x in np.asarray(values, dtype=np.float64).tolist():
                s += x
            mean = s / len(values)
        return KrigingMethod.Simple(mean)
    if method == "universal":
        return KrigingMethod.Universal(degree)
    raise InvalidParameter(f"unknown method '{method}' (expected ordinary, simple or universal)")


def _search(search_azimuth, search_ratio, search_ratio_z, search_dip, search_rake):
    """anisotropic_search (python/src/lib.rs:28-42)."""
    if search_azimuth is None:
        return None
    return Anisotropy(search_azimuth, search_ratio, search_ratio_z, search_dip, search_rake)


def krige(x, y, values, model: VariogramModel, target_x, target_y, method="ordinary",
          mean=None, degree=1, max_neighbors=None, radius=None, min_neighbors=None, octant=None,
          measurement_error=None, search_azimuth=None, search_ratio=1.0, search_ratio_z=1.0,
          search_dip=0.0, search_rake=0.0) -> ArrayPair:
    """krige (python/src/lib.rs:479-555)."""
    if len(target_x) != len(target_y):
        raise ValueError("target_x and target_y differ in length")
    c, v = _coords2(x, y, values)
    cfg = KrigingConfig(method=_method(method, mean, degree, v), max_neighbors=max_neighbors,
                        search_radius=radius, min_neighbors=min_neighbors,
                        max_per_octant=octant,
                        anisotropic_search=_search(search_azimuth, search_ratio, search_ratio_z,
                                                   search_dip, search_rake))
    t = np.column_stack([np.asarray(target_x, dtype=np.float64),
                         np.asarray(target_y, dtype=np.float64)])
    if measurement_error is not None:
        k = Kriging.with_measurement_error(c, v, model, cfg, measurement_error)
    else:
        k = Kriging(c, v, model, cfg)
    val, var, _, _ = k.predict_many(t)
    return val, var


def krige_grid(x, y, values, model: VariogramModel, bbox, nx, ny, method="ordinary", mean=None,
               degree=1, max_neighbors=None, radius=None, block=None, block_discr=(4, 4),
               min_neighbors=None, octant=None, search_azimuth=None, search_ratio=1.0,
               search_ratio_z=1.0, search_dip=0.0, search_rake=0.0) -> ArrayPair:
    """krige_grid (python/src/lib.rs:622-689): row-major cell centres, y
    increasing; block kriging when `block = (width, height)`."""
    c, v = _coords2(x, y, values)
    grid = Grid2D.from_bbox([bbox[0], bbox[1]], [bbox[2], bbox[3]], nx, ny)
    cfg = KrigingConfig(method=_method(method, mean, degree, v), max_neighbors=max_neighbors,
                        search_radius=radius, min_neighbors=min_neighbors,
                        max_per_octant=octant,
                        anisotropic_search=_search(search_azimuth, search_ratio, search_ratio_z,
                                                   search_dip, search_rake))
    k = Kriging(c, v, model, cfg)
    if block is not None:
        return k.predict_block_grid(grid, [block[0], block[1]], [block_discr[0], block_discr[1]])
    return k.predict_grid(grid)


def krige_3d(x, y, z, values, model: VariogramModel, target_x, target_y, target_z,
             method="ordinary", mean=None, degree=1, max_neighbors=None,
             radius=None) -> ArrayPair:
    """krige_3d (python/src/lib.rs:1647-1706)."""
    if not (len(target_x) == len(target_y) == len(target_z)):
        raise ValueError("target_x, target_y and target_z differ in length")
    x, y, z = (np.asarray(a, dtype=np.float64) for a in (x, y, z))
    v = np.asarray(values, dtype=np.float64)
    if not (len(x) == len(y) == len(z) == len(v)):
        from . import DimensionMismatch
        raise DimensionMismatch("x, y, z and values differ in length")
    cfg = KrigingConfig(method=_method(method, mean, degree, v), max_neighbors=max_neighbors,
                        search_radius=radius)
    t = np.column_stack([np.asarray(a, dtype=np.float64) for a in (target_x, target_y, target_z)])
    val, var, _, _ = Kriging(np.column_stack([x, y, z]), v, model, cfg).predict_many(t)
    return val, var


def vecchia_loglik(x, y, values, model: VariogramModel, m=20) -> float:
    """vecchia_loglik (python/src/lib.rs:476-490)."""
    c, v = _coords2(x, y, values)
    return _vecchia_loglik(c, v, model, m)


def loo_cv(x, y, values, model: VariogramModel, method="ordinary", mean=None, degree=1,
           max_neighbors=None, radius=None, folds=None, seed=0, min_neighbors=None,
           octant=None, blocks=None) -> dict:
    """loo_cv (python/src/lib.rs:691-752): leave-one-out, k-fold (`folds`) or
    spatial block CV (`blocks`); a dict of the accuracy measures and arrays."""
    c, v = _coords2(x, y, values)
    cfg = KrigingConfig(method=_method(method, mean, degree, v), max_neighbors=max_neighbors,
                        search_radius=radius, min_neighbors=min_neighbors, max_per_octant=octant)
    if blocks is not None and folds is not None:
        raise ValueError("blocks and folds are mutually exclusive")
    if blocks is not None:
        cv = _cv.block_cv(c, v, model, cfg, [blocks[0], blocks[1]])
    elif folds is not None:
        cv = _cv.k_fold(c, v, model, cfg, folds, seed)
    else:
        cv = _cv.leave_one_out(c, v, model, cfg)
    return {"me": cv.mean_error(), "mae": cv.mae(), "mse": cv.mse(), "rmse": cv.rmse(),
            "msdr": cv.msdr(), "rme": cv.rme(), "rmae": cv.rmae(), "rrmse": cv.rrmse(),
            "vecv": cv.vecv(), "e1": cv.e1(), "observed": cv.observed,
            "predicted": cv.predicted, "variance": cv.variance}
