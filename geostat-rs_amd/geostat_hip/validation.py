"""Cross-validation over the GPU predictor (validation.rs:10-400).

The folds run in the library (gs_cv_*): each fold builds a device-resident
predictor on the points outside it and predicts its members.  CvResult's
accuracy measures are plain host arithmetic over the three arrays, in the
reference's formulas and summation order.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _abi
from . import (Context, DimensionMismatch, KrigingConfig, VariogramModel, _check, _f64)

_EPS = 2.220446049250313e-16   # f64::EPSILON


def _seq_sum(xs) -> float:
    s = -0.0
    for x in xs:
        s += float(x)
    return s


@dataclass
class CvResult:
    """CvResult (validation.rs:10-130): observed, predicted and kriging
    variance per point, in dataset order."""
    observed: np.ndarray
    predicted: np.ndarray
    variance: np.ndarray

    def _res(self):
        return [p - o for p, o in zip(self.predicted.tolist(), self.observed.tolist())]

    def mean_error(self) -> float:
        return _seq_sum(self._res()) / len(self.observed)

    def mae(self) -> float:
        return _seq_sum(abs(e) for e in self._res()) / len(self.observed)

    def mse(self) -> float:
        return _seq_sum(e * e for e in self._res()) / len(self.observed)

    def rmse(self) -> float:
        return math.sqrt(self.mse())

    def _obs_mean(self) -> float:
        return _seq_sum(self.observed.tolist()) / len(self.observed)

    def rme(self) -> float:
        m = self._obs_mean()
        return math.nan if abs(m) < _EPS else (-self.mean_error()) / m * 100.0

    def rmae(self) -> float:
        m = self._obs_mean()
        return math.nan if abs(m) < _EPS else self.mae() / m * 100.0

    def rrmse(self) -> float:
        m = self._obs_mean()
        return math.nan if abs(m) < _EPS else self.rmse() / m * 100.0

    def vecv(self) -> float:
        m = self._obs_mean()
        sse = _seq_sum(e * e for e in self._res())
        sst = _seq_sum((o - m) * (o - m) for o in self.observed.tolist())
        return math.nan if sst <= 0.0 else (1.0 - sse / sst) * 100.0

    def e1(self) -> float:
        m = self._obs_mean()
        sae = _seq_sum(abs(e) for e in self._res())
        sad = _seq_sum(abs(o - m) for o in self.observed.tolist())
        return math.nan if sad <= 0.0 else (1.0 - sae / sad) * 100.0

    def msdr(self) -> float:
        s, k = 0.0, 0
        for p, o, v in zip(self.predicted.tolist(), self.observed.tolist(),
                           self.variance.tolist()):
            if v > 1e-12:
                e = p - o
                s += e * e / v
                k += 1
        return math.nan if k == 0 else s / k


def _prep(coords, values, ctx):
    coords = _f64(coords)
    values = _f64(values).reshape(-1)
    if coords.ndim != 2 or coords.shape[1] not in (2, 3):
        raise DimensionMismatch(f"coords must be (n, 2|3), got {coords.shape}")
    if coords.shape[0] != values.shape[0]:
        raise DimensionMismatch(f"{coords.shape[0]} coordinates vs {values.shape[0]} values")
    return coords, values, ctx or Context.default()


def _run(fn, coords, values, model: VariogramModel, config: KrigingConfig, ctx, *extra):
    coords, values, ctx = _prep(coords, values, ctx)
    n = coords.shape[0]
    pred, var = np.empty(n), np.empty(n)
    cm, cc = model._c(), config._c()
    _check(fn(ctx._h, coords.shape[1], _abi.dptr(coords), _abi.dptr(values), n, C.byref(cm),
              C.byref(cc), *extra, _abi.dptr(pred), _abi.dptr(var)))
    return CvResult(values.copy(), pred, var)


def leave_one_out(coords, values, model: VariogramModel, config: KrigingConfig,
                  ctx: Optional[Context] = None) -> CvResult:
    """leave_one_out (validation.rs:136-159)."""
    return _run(_abi.lib().gs_cv_leave_one_out, coords, values, model, config, ctx)


def leave_one_out_with_drift(coords, values, drift_data, model: VariogramModel,
                             config: KrigingConfig, ctx: Optional[Context] = None) -> CvResult:
    """leave_one_out_with_drift (validation.rs:163-194)."""
    d = _f64(drift_data)
    n = _f64(values).reshape(-1).shape[0]
    if d.ndim != 2 or d.shape[0] != n:
        raise DimensionMismatch(f"{d.shape[0] if d.ndim else 0} drift rows vs {n} data points")
    return _run(_abi.lib().gs_cv_leave_one_out_with_drift, coords, values, model, config, ctx,
                d.shape[1], _abi.dptr(d))


def k_fold(coords, values, model: VariogramModel, config: KrigingConfig, k: int, seed: int,
           ctx: Optional[Context] = None) -> CvResult:
    """k_fold (validation.rs:272-292)."""
    return _run(_abi.lib().gs_cv_k_fold, coords, values, model, config, ctx, int(k),
                int(seed) & 0xFFFFFFFFFFFFFFFF)


def block_cv(coords, values, model: VariogramModel, config: KrigingConfig,
             blocks_per_dim: Sequence[int], ctx: Optional[Context] = None) -> CvResult:
    """block_cv (validation.rs:348-373)."""
    b = np.ascontiguousarray(np.asarray(blocks_per_dim, dtype=np.int64))
    dim = _f64(coords).shape[1]
    if b.shape[0] != dim:
        raise DimensionMismatch(f"{b.shape[0]} block counts for {dim} dimensions")
    return _run(_abi.lib().gs_cv_block, coords, values, model, config, ctx, _abi.i64ptr(b))
