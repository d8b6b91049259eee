"""Seeded synthetic workloads for the benchmark configurations.

Data follow the reference's bench recipe (benches/engine.rs:12-23): per
point x = U*L, y = U*L [, z = U*Lz], v = sin(x/150) + cos(y/200)
[+ 0.001 z] + 0.3 N, drawn from the xoshiro256++ stream of rng.rs.

  ok2d_small   n = 2,000 on [0,1000]^2, spherical (nugget 0.1, sill 1,
               range 250), 20 nearest, 100x100 grid (10^4 targets)
  ok2d_large   n = 200,000 on [0,10^4]^2, Matern 3/2 (nugget 0.05, sill 1,
               range 800, azimuth 30, ratio 0.5), 32 nearest,
               2000x2000 grid (4*10^6 targets)
  uk3d_octant  n = 10^6 in [0,5000]^2 x [0,500], linear drift in z,
               nugget 0.1 + exponential(600) + spherical(1500) rotated,
               universal kriging degree 1, octant search (8 per octant,
               min 8) in a rotated ellipsoid, 200x200x100 grid
  vecchia      n = 10^6 on [0,1]^2, Matern 5/2 (range 0.05, sill 1,
               nugget 0.01) field by Vecchia simulation (maxmin, m = 30),
               64 (range, sill, nugget) vectors log-uniform within +-1 decade
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _abi
from . import (Anisotropy, Grid2D, Grid3D, KrigingConfig, KrigingMethod, ModelKind, Structure,
               VariogramModel)

SEED_BASE = 42


def points(seed: int, n: int, ext, recipe: int = 0, noise: float = 0.3):
    """(coords (n, D), values (n,)) from the seeded generator."""
    ext = np.ascontiguousarray(np.asarray(ext, dtype=np.float64))
    dim = ext.shape[0]
    coords = np.empty((n, dim))
    values = np.empty(n)
    _abi.synth().gs_synth_points(seed, n, dim, _abi.dptr(ext), recipe, noise, _abi.dptr(coords),
                                 _abi.dptr(values))
    return coords, values


def raw_u64(seed: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.uint64)
    _abi.synth().gs_synth_raw(seed, count, out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return out


def uniforms(seed: int, count: int) -> np.ndarray:
    out = np.empty(count)
    _abi.synth().gs_synth_uniform(seed, count, _abi.dptr(out))
    return out


@dataclass
class KrigingWorkload:
    name: str
    coords: np.ndarray
    values: np.ndarray
    model: VariogramModel
    config: KrigingConfig
    grid: object

    def targets(self) -> np.ndarray:
        return self.grid.centers()


def ok2d_small() -> KrigingWorkload:
    c, v = points(SEED_BASE + 0, 2000, [1000.0, 1000.0])
    model = VariogramModel(0.1, [Structure(ModelKind.SPHERICAL, 1.0, 250.0)])
    cfg = KrigingConfig(max_neighbors=20)
    grid = Grid2D.from_bbox([0.0, 0.0], [1000.0, 1000.0], 100, 100)
    return KrigingWorkload("ok2d_small", c, v, model, cfg, grid)


def ok2d_large(n: int = 200_000, nx: int = 2000, ny: int = 2000) -> KrigingWorkload:
    c, v = points(SEED_BASE + 1, n, [10_000.0, 10_000.0])
    model = VariogramModel(0.05, [Structure(ModelKind.MATERN15, 1.0, 800.0,
                                            Anisotropy(30.0, 0.5))])
    cfg = KrigingConfig(max_neighbors=32)
    grid = Grid2D.from_bbox([0.0, 0.0], [10_000.0, 10_000.0], nx, ny)
    return KrigingWorkload("ok2d_large", c, v, model, cfg, grid)


UK3D_ANIS = Anisotropy(azimuth_deg=45.0, ratio=0.6, ratio_z=0.2, dip_deg=10.0, rake_deg=5.0)
UK3D_RADIUS = 230.0


def uk3d_octant(n: int = 1_000_000, nx: int = 200, ny: int = 200, nz: int = 100) -> KrigingWorkload:
    c, v = points(SEED_BASE + 2, n, [5000.0, 5000.0, 500.0], recipe=1)
    model = VariogramModel(0.1, [
        Structure(ModelKind.EXPONENTIAL, 0.6, 600.0, UK3D_ANIS),
        Structure(ModelKind.SPHERICAL, 0.4, 1500.0, UK3D_ANIS),
    ])
    cfg = KrigingConfig(method=KrigingMethod.Universal(1), max_neighbors=64, min_neighbors=8,
                        max_per_octant=8, search_radius=UK3D_RADIUS, anisotropic_search=UK3D_ANIS)
    grid = Grid3D.from_bbox([0.0, 0.0, 0.0], [5000.0, 5000.0, 500.0], nx, ny, nz)
    return KrigingWorkload("uk3d_octant", c, v, model, cfg, grid)


def vecchia_models(count: int = 64, seed: int = SEED_BASE + 30):
    """`count` Matern-5/2 models with (range, sill, nugget) log-uniform
    within +-1 decade of (0.05, 1, 0.01)."""
    u = uniforms(seed, 3 * count).reshape(count, 3)
    out = []
    for a, b, c in u:
        rng = 0.05 * 10.0 ** (2.0 * a - 1.0)
        sill = 1.0 * 10.0 ** (2.0 * b - 1.0)
        nug = 0.01 * 10.0 ** (2.0 * c - 1.0)
        out.append(VariogramModel(nug, [Structure(ModelKind.MATERN25, sill, rng)]))
    return out


def vecchia_field(coords, order, nb, nb_count, m, nugget=0.01, sill=1.0, rng=0.05,
                  seed: int = SEED_BASE + 3, threads: int = 8) -> np.ndarray:
    """Values of a zero-mean Matern-5/2 field drawn by sequential Vecchia
    simulation over the plan (order, nb, nb_count)."""
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    n, dim = coords.shape
    vals = np.zeros(n)
    _abi.synth().gs_synth_vecchia_field(seed, n, dim, _abi.dptr(coords),
                                        _abi.i64ptr(np.ascontiguousarray(order, dtype=np.int64)),
                                        _abi.i32ptr(np.ascontiguousarray(nb, dtype=np.int32)),
                                        _abi.i32ptr(np.ascontiguousarray(nb_count, dtype=np.int32)),
                                        m, nugget, sill, rng, threads, _abi.dptr(vals))
    return vals


def vecchia_c3(n: int = 1_000_000, seed: int = SEED_BASE + 3):
    """Coordinates of the Vecchia likelihood workload: n points uniform on
    [0, 1]^2 (values are drawn over the plan by vecchia_field)."""
    c, _ = points(seed, n, [1.0, 1.0], recipe=2)
    return c


@dataclass
class ServingRequest:
    arrival_gap: float      # in units of the mean gap
    n: int
    targets: int
    k: int
    kind: int
    seed: int


SERVING_KINDS = (ModelKind.SPHERICAL, ModelKind.EXPONENTIAL, ModelKind.MATERN15,
                 ModelKind.MATERN25)


def serving_requests(count: int = 40, seed: int = SEED_BASE + 4, n_max: int = 200_000,
                     t_max: int = 1_000_000):
    """A request mix: exponential inter-arrival gaps (Poisson arrivals),
    n log-uniform in [10^3, n_max], targets log-uniform in [10^4, t_max],
    k in {16, 32, 64}, model kind in {Sph, Exp, Mat15, Mat25}."""
    u = uniforms(seed, 5 * count).reshape(count, 5)
    out = []
    for i, (a, b, c, d, e) in enumerate(u):
        out.append(ServingRequest(
            arrival_gap=-np.log(max(1.0 - a, 1e-300)),
            n=int(round(10.0 ** (3.0 + b * (np.log10(n_max) - 3.0)))),
            targets=int(round(10.0 ** (4.0 + c * (np.log10(t_max) - 4.0)))),
            k=(16, 32, 64)[min(int(d * 3), 2)],
            kind=SERVING_KINDS[min(int(e * 4), 3)],
            seed=seed + 1 + i))
    return out


def serving_case(req: ServingRequest, ext: float = 10_000.0):
    """(coords, values, model, config, targets) of one serving request:
    data by the reference bench recipe on [0, ext]^2, targets uniform in
    the same square, OK with max_neighbors = k."""
    c, v = points(req.seed, req.n, [ext, ext], recipe=0)
    t = uniforms(req.seed + 100_000, 2 * req.targets).reshape(req.targets, 2) * ext
    model = VariogramModel(0.05, [Structure(req.kind, 1.0, 800.0)])
    cfg = KrigingConfig(max_neighbors=req.k)
    return c, v, model, cfg, t
