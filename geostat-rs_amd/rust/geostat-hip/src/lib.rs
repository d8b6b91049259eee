//! geostat-hip: the moving-neighbourhood kriging predictor and the Vecchia
//! log-likelihood of geostat-core on one MI355X, behind geostat-core's own
//! types.  A caller switches by building `geostat_hip::Kriging` where it
//! built `geostat_core::Kriging` (kriging.rs:166) and calling the same
//! methods; results match the rayon path (neighbour lists bit-exact,
//! estimates within 1e-10 in FP64), per-target failures are NaN estimates
//! as in predict_many (kriging.rs:648-652), build errors are the same
//! GeostatError variants.
mod ffi;

use std::ffi::{c_void, CStr};
use std::ptr;

use geostat_core::{
    block_offsets, Anisotropy, CvResult, GeostatError, Grid2D, KrigingConfig, KrigingEstimate,
    KrigingMethod, ModelKind, PointSet, Result, Structure, VariogramModel, VecchiaFit,
};

pub use ffi::{GS_ABI_VERSION, GS_PREC_FP32_REFINED, GS_PREC_FP64};

/// Arithmetic of the covariance build and solve.
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub enum Precision {
    /// FP64 throughout (the reference's arithmetic).
    Fp64,
    /// FP32 factorization with an FP64 correction (within 1e-6 of FP64).
    Fp32Refined,
}

impl Precision {
    fn code(self) -> i32 {
        match self {
            Precision::Fp64 => ffi::GS_PREC_FP64,
            Precision::Fp32Refined => ffi::GS_PREC_FP32_REFINED,
        }
    }
}

/// The calling thread's last error message from the library.
pub fn last_error() -> String {
    unsafe {
        let p = ffi::gs_last_error();
        if p.is_null() {
            String::new()
        } else {
            CStr::from_ptr(p).to_string_lossy().into_owned()
        }
    }
}

/// gs_status -> GeostatError (error.rs:15-40; codes follow its declaration
/// order).  Device failures surface as InvalidParameter with the HIP text.
fn status_to_result(st: i32, dup: [i64; 2]) -> Result<()> {
    match st {
        ffi::GS_OK => Ok(()),
        ffi::GS_ERR_INSUFFICIENT_DATA => Err(GeostatError::InsufficientData(last_error())),
        ffi::GS_ERR_DIMENSION_MISMATCH => Err(GeostatError::DimensionMismatch(last_error())),
        ffi::GS_ERR_SINGULAR_SYSTEM => Err(GeostatError::SingularSystem(last_error())),
        ffi::GS_ERR_INVALID_PARAMETER => Err(GeostatError::InvalidParameter(last_error())),
        ffi::GS_ERR_NO_NEIGHBORS => Err(GeostatError::NoNeighbors),
        ffi::GS_ERR_DUPLICATE_POINTS => {
            Err(GeostatError::DuplicatePoints(dup[0] as usize, dup[1] as usize))
        }
        _ => Err(GeostatError::InvalidParameter(format!("device failure: {}", last_error()))),
    }
}

/// The error a single-target call returns for a per-target status.
fn target_error(st: i32) -> GeostatError {
    match st {
        ffi::GS_ERR_INSUFFICIENT_DATA => {
            GeostatError::InsufficientData("fewer neighbours than drift terms".into())
        }
        ffi::GS_ERR_SINGULAR_SYSTEM => GeostatError::SingularSystem("kriging system".into()),
        ffi::GS_ERR_NO_NEIGHBORS => GeostatError::NoNeighbors,
        _ => GeostatError::InvalidParameter(format!("target failed with status {st}")),
    }
}

fn to_gs_anis(a: &Anisotropy) -> ffi::GsAnisotropy {
    ffi::GsAnisotropy {
        azimuth_deg: a.azimuth_deg,
        ratio: a.ratio,
        ratio_z: a.ratio_z,
        dip_deg: a.dip_deg,
        rake_deg: a.rake_deg,
    }
}

/// ModelKind (model.rs:47-95) -> (gs_kind, parameter).
fn kind_code(k: ModelKind) -> (i32, f64) {
    match k {
        ModelKind::Spherical => (0, 0.0),
        ModelKind::Exponential => (1, 0.0),
        ModelKind::Gaussian => (2, 0.0),
        ModelKind::Matern15 => (3, 0.0),
        ModelKind::Matern25 => (4, 0.0),
        ModelKind::Matern(nu) => (5, nu),
        ModelKind::Circular => (6, 0.0),
        ModelKind::Stable(alpha) => (7, alpha),
        ModelKind::Hole => (8, 0.0),
        ModelKind::Wave => (9, 0.0),
        ModelKind::Power(theta) => (10, theta),
    }
}

fn from_kind_code(code: i32, param: f64) -> ModelKind {
    match code {
        0 => ModelKind::Spherical,
        1 => ModelKind::Exponential,
        2 => ModelKind::Gaussian,
        3 => ModelKind::Matern15,
        4 => ModelKind::Matern25,
        5 => ModelKind::Matern(param),
        6 => ModelKind::Circular,
        7 => ModelKind::Stable(param),
        8 => ModelKind::Hole,
        9 => ModelKind::Wave,
        _ => ModelKind::Power(param),
    }
}

/// VariogramModel (model.rs:543-549) as the ABI's fixed-size model.
fn to_gs_model(m: &VariogramModel) -> Result<ffi::GsModel> {
    if m.structures.len() > ffi::GS_MAX_STRUCTURES {
        return Err(GeostatError::InvalidParameter(format!(
            "unsupported on GPU: {} nested structures (at most {})",
            m.structures.len(),
            ffi::GS_MAX_STRUCTURES
        )));
    }
    let mut g = ffi::GsModel { nugget: m.nugget, n_structures: m.structures.len() as i32,
                               ..Default::default() };
    for (i, s) in m.structures.iter().enumerate() {
        let (kind, param) = kind_code(s.kind);
        g.structures[i] = ffi::GsStructure {
            kind,
            param,
            sill: s.sill,
            range: s.range,
            has_anis: s.anis.is_some() as i32,
            anis: s.anis.as_ref().map(to_gs_anis).unwrap_or_default(),
        };
    }
    Ok(g)
}

fn from_gs_model(g: &ffi::GsModel) -> Result<VariogramModel> {
    let structures = g.structures[..g.n_structures as usize]
        .iter()
        .map(|s| Structure::new(from_kind_code(s.kind, s.param), s.sill, s.range))
        .collect();
    VariogramModel::new(g.nugget, structures)
}

fn opt(v: Option<usize>) -> i64 {
    v.map(|x| x as i64).unwrap_or(ffi::GS_NONE)
}

/// KrigingConfig (kriging.rs:73-115) with None as GS_NONE.
fn to_gs_config(c: &KrigingConfig) -> ffi::GsKConfig {
    let (method, simple_mean, degree) = match c.method {
        KrigingMethod::Simple { mean } => (ffi::GS_SIMPLE, mean, 0),
        KrigingMethod::Ordinary => (ffi::GS_ORDINARY, 0.0, 0),
        KrigingMethod::Universal { degree } => (ffi::GS_UNIVERSAL, 0.0, degree as i32),
        KrigingMethod::ExternalDrift { n_vars } => (ffi::GS_EXTERNAL_DRIFT, 0.0, n_vars as i32),
    };
    ffi::GsKConfig {
        method,
        simple_mean,
        degree,
        max_neighbors: opt(c.max_neighbors),
        min_neighbors: opt(c.min_neighbors),
        max_per_octant: opt(c.max_per_octant),
        has_search_radius: c.search_radius.is_some() as i32,
        search_radius: c.search_radius.unwrap_or(0.0),
        has_anisotropic_search: c.anisotropic_search.is_some() as i32,
        anisotropic_search: c.anisotropic_search.as_ref().map(to_gs_anis).unwrap_or_default(),
    }
}

fn flat<const D: usize>(points: &[[f64; D]]) -> &[f64] {
    // [[f64; D]] is laid out as D * len contiguous f64
    unsafe { std::slice::from_raw_parts(points.as_ptr() as *const f64, points.len() * D) }
}

fn rows(v: &[Vec<f64>], width: usize) -> Result<Vec<f64>> {
    let mut out = Vec::with_capacity(v.len() * width);
    for r in v {
        if r.len() != width {
            return Err(GeostatError::DimensionMismatch(format!(
                "drift row has {} values, expected {width}",
                r.len()
            )));
        }
        out.extend_from_slice(r);
    }
    Ok(out)
}

/// A device and its default stream (no reference counterpart: the rayon
/// path has no device state).
pub struct Context {
    h: *mut ffi::gs_ctx,
}

// The library serialises use of a context's shared state internally.
unsafe impl Send for Context {}
unsafe impl Sync for Context {}

impl Context {
    pub fn new(device: i32) -> Result<Self> {
        let mut h = ptr::null_mut();
        status_to_result(unsafe { ffi::gs_ctx_create(device, &mut h) }, [0, 0])?;
        Ok(Self { h })
    }

    pub fn synchronize(&self) -> Result<()> {
        status_to_result(unsafe { ffi::gs_ctx_synchronize(self.h) }, [0, 0])
    }
}

impl Drop for Context {
    fn drop(&mut self) {
        unsafe { ffi::gs_ctx_destroy(self.h) }
    }
}

/// Kriging::<D, VariogramModel> (kriging.rs:142-162) with data, search
/// index and model resident on the GPU.
pub struct Kriging<const D: usize> {
    h: *mut ffi::gs_kriging,
    method: KrigingMethod,
}

unsafe impl<const D: usize> Send for Kriging<D> {}
unsafe impl<const D: usize> Sync for Kriging<D> {}

impl<const D: usize> Kriging<D> {
    /// Kriging::new (kriging.rs:166): the same validation and errors.
    pub fn new(ctx: &Context, data: &PointSet<D>, model: &VariogramModel,
               config: KrigingConfig) -> Result<Self> {
        let (m, c) = (to_gs_model(model)?, to_gs_config(&config));
        let mut h = ptr::null_mut();
        let mut dup = [0i64; 2];
        let st = unsafe {
            ffi::gs_kriging_new(ctx.h, D as i32, flat(data.coords()).as_ptr(),
                                data.values().as_ptr(), data.len() as i64, &m, &c, &mut h,
                                dup.as_mut_ptr())
        };
        status_to_result(st, dup)?;
        Ok(Self { h, method: config.method })
    }

    /// Kriging::with_measurement_error (kriging.rs:181-218).
    pub fn with_measurement_error(ctx: &Context, data: &PointSet<D>, model: &VariogramModel,
                                  config: KrigingConfig, errors: &[f64]) -> Result<Self> {
        if errors.len() != data.len() {
            return Err(GeostatError::DimensionMismatch(format!(
                "{} error variances vs {} data points", errors.len(), data.len())));
        }
        let (m, c) = (to_gs_model(model)?, to_gs_config(&config));
        let mut h = ptr::null_mut();
        let mut dup = [0i64; 2];
        let st = unsafe {
            ffi::gs_kriging_new_with_measurement_error(
                ctx.h, D as i32, flat(data.coords()).as_ptr(), data.values().as_ptr(),
                data.len() as i64, &m, &c, errors.as_ptr(), &mut h, dup.as_mut_ptr())
        };
        status_to_result(st, dup)?;
        Ok(Self { h, method: config.method })
    }

    /// Kriging::with_external_drift (kriging.rs:223-260): one row of
    /// covariates per datum.
    pub fn with_external_drift(ctx: &Context, data: &PointSet<D>, model: &VariogramModel,
                               config: KrigingConfig, drift_data: &[Vec<f64>]) -> Result<Self> {
        let n_vars = match config.method {
            KrigingMethod::ExternalDrift { n_vars } => n_vars,
            _ => drift_data.first().map(|r| r.len()).unwrap_or(0),
        };
        if drift_data.len() != data.len() {
            return Err(GeostatError::DimensionMismatch(format!(
                "{} drift rows vs {} data points", drift_data.len(), data.len())));
        }
        let d = rows(drift_data, n_vars)?;
        let (m, c) = (to_gs_model(model)?, to_gs_config(&config));
        let mut h = ptr::null_mut();
        let mut dup = [0i64; 2];
        let st = unsafe {
            ffi::gs_kriging_new_with_external_drift(
                ctx.h, D as i32, flat(data.coords()).as_ptr(), data.values().as_ptr(),
                data.len() as i64, &m, &c, n_vars as i32, d.as_ptr(), &mut h, dup.as_mut_ptr())
        };
        status_to_result(st, dup)?;
        Ok(Self { h, method: config.method })
    }

    fn estimates(&self, v: &[f64], var: &[f64], lag: &[f64], st: &[i32]) -> Vec<KrigingEstimate> {
        let simple = matches!(self.method, KrigingMethod::Simple { .. });
        (0..v.len())
            .map(|i| KrigingEstimate {
                value: v[i],
                variance: var[i],
                lagrange: (!simple && st[i] == 0).then_some(lag[i]),
            })
            .collect()
    }

    /// Kriging::predict_many (kriging.rs:648-652) at `precision`: failed
    /// targets are NaN estimates.  Panics only on a device failure.
    pub fn predict_many_with(&self, targets: &[[f64; D]], precision: Precision)
                             -> Vec<KrigingEstimate> {
        let t = targets.len();
        let (mut v, mut var, mut lag) = (vec![0.0; t], vec![0.0; t], vec![0.0; t]);
        let mut st = vec![0i32; t];
        if t > 0 {
            let rc = unsafe {
                ffi::gs_kriging_predict_many(self.h, flat(targets).as_ptr(), t as i64,
                                             precision.code(), v.as_mut_ptr(), var.as_mut_ptr(),
                                             lag.as_mut_ptr(), st.as_mut_ptr())
            };
            assert_eq!(rc, ffi::GS_OK, "device failure: {}", last_error());
        }
        self.estimates(&v, &var, &lag, &st)
    }

    /// Kriging::predict_many (kriging.rs:648-652) in FP64.
    pub fn predict_many(&self, targets: &[[f64; D]]) -> Vec<KrigingEstimate> {
        self.predict_many_with(targets, Precision::Fp64)
    }

    /// Kriging::predict (kriging.rs:494-501): the per-target error.
    pub fn predict(&self, target: [f64; D]) -> Result<KrigingEstimate> {
        let t = [target];
        let (mut v, mut var, mut lag, mut st) = ([0.0], [0.0], [0.0], [0i32]);
        let rc = unsafe {
            ffi::gs_kriging_predict_many(self.h, flat(&t).as_ptr(), 1, ffi::GS_PREC_FP64,
                                         v.as_mut_ptr(), var.as_mut_ptr(), lag.as_mut_ptr(),
                                         st.as_mut_ptr())
        };
        status_to_result(rc, [0, 0])?;
        if st[0] != 0 {
            return Err(target_error(st[0]));
        }
        Ok(self.estimates(&v, &var, &lag, &st)[0])
    }

    /// Kriging::predict_many_with_drift (kriging.rs:656-672).
    pub fn predict_many_with_drift(&self, targets: &[[f64; D]], drifts: &[Vec<f64>])
                                   -> Vec<KrigingEstimate> {
        let t = targets.len();
        let width = drifts.first().map(|r| r.len()).unwrap_or(0);
        let d = match rows(drifts, width) {
            Ok(d) if drifts.len() == t => d,
            _ => return vec![KrigingEstimate { value: f64::NAN, variance: f64::NAN,
                                               lagrange: None }; t],
        };
        let (mut v, mut var, mut lag) = (vec![0.0; t], vec![0.0; t], vec![0.0; t]);
        let mut st = vec![0i32; t];
        if t > 0 {
            let rc = unsafe {
                ffi::gs_kriging_predict_many_with_drift(
                    self.h, flat(targets).as_ptr(), t as i64, d.as_ptr(), ffi::GS_PREC_FP64,
                    v.as_mut_ptr(), var.as_mut_ptr(), lag.as_mut_ptr(), st.as_mut_ptr())
            };
            assert_eq!(rc, ffi::GS_OK, "device failure: {}", last_error());
        }
        self.estimates(&v, &var, &lag, &st)
    }

    /// Kriging::predict_with_drift (kriging.rs:505-522).
    pub fn predict_with_drift(&self, target: [f64; D], target_drift: &[f64])
                              -> Result<KrigingEstimate> {
        let e = self.predict_many_with_drift(&[target], &[target_drift.to_vec()])[0];
        if e.value.is_nan() {
            return Err(GeostatError::InvalidParameter("target failed".into()));
        }
        Ok(e)
    }

    /// Kriging::predict_block (kriging.rs:678-777) at each centre.
    pub fn predict_block_many(&self, centers: &[[f64; D]], offsets: &[[f64; D]])
                              -> Vec<KrigingEstimate> {
        let t = centers.len();
        let (mut v, mut var, mut lag) = (vec![0.0; t], vec![0.0; t], vec![0.0; t]);
        let mut st = vec![0i32; t];
        if t > 0 {
            let rc = unsafe {
                ffi::gs_kriging_predict_block_many(
                    self.h, flat(centers).as_ptr(), t as i64,
                    if offsets.is_empty() { ptr::null() } else { flat(offsets).as_ptr() },
                    offsets.len() as i32, ffi::GS_PREC_FP64, v.as_mut_ptr(), var.as_mut_ptr(),
                    lag.as_mut_ptr(), st.as_mut_ptr())
            };
            assert_eq!(rc, ffi::GS_OK, "device failure: {}", last_error());
        }
        self.estimates(&v, &var, &lag, &st)
    }

    /// Kriging::predict_block (kriging.rs:678-777).
    pub fn predict_block(&self, center: [f64; D], offsets: &[[f64; D]]) -> Result<KrigingEstimate> {
        let e = self.predict_block_many(&[center], offsets)[0];
        if e.value.is_nan() {
            return Err(GeostatError::InvalidParameter("block failed".into()));
        }
        Ok(e)
    }

    /// Kriging::neighbors (kriging.rs:441-490): original data indices in
    /// the reference's order.
    pub fn neighbors(&self, target: [f64; D]) -> Result<Vec<usize>> {
        let max_out = unsafe { ffi::gs_kriging_max_neighbors(self.h) }.max(1);
        let mut idx = vec![-1i32; max_out as usize];
        let mut cnt = [0i32];
        let t = [target];
        let rc = unsafe {
            ffi::gs_kriging_neighbors(self.h, flat(&t).as_ptr(), 1, max_out, idx.as_mut_ptr(),
                                      cnt.as_mut_ptr())
        };
        status_to_result(rc, [0, 0])?;
        Ok(idx.iter().take(cnt[0] as usize).map(|&i| i as usize).collect())
    }

    /// predict_many on device-resident buffers, enqueued on `stream` (a
    /// hipStream_t; null is the null stream).
    ///
    /// # Safety
    /// Every pointer must be a device allocation on this handle's GPU:
    /// `targets` with t * D doubles, `value` / `variance` / `lagrange` with
    /// t doubles, `status` with t int32 (the last two may be null).
    pub unsafe fn predict_many_device(&self, targets: *const f64, t: usize, precision: Precision,
                                      value: *mut f64, variance: *mut f64, lagrange: *mut f64,
                                      status: *mut i32, stream: *mut c_void) -> Result<()> {
        status_to_result(ffi::gs_kriging_predict_many_device(self.h, targets, t as i64,
                                                             precision.code(), value, variance,
                                                             lagrange, status, stream),
                         [0, 0])
    }

    /// predict_many_device captured as a HIP graph over these buffers: the
    /// call runs once, then every `PredictGraph::launch` repeats it on
    /// whatever `targets` holds.
    ///
    /// # Safety
    /// As `predict_many_device`; the buffers must outlive the graph.
    pub unsafe fn capture_graph(&self, targets: *const f64, t: usize, precision: Precision,
                                value: *mut f64, variance: *mut f64, lagrange: *mut f64,
                                status: *mut i32) -> Result<PredictGraph<'_, D>> {
        let mut g = std::ptr::null_mut();
        status_to_result(ffi::gs_kriging_graph_create(self.h, targets, t as i64, precision.code(),
                                                      value, variance, lagrange, status, &mut g),
                         [0, 0])?;
        Ok(PredictGraph { g, _k: self })
    }
}

/// A captured predict call; borrows its predictor so that it cannot outlive it.
pub struct PredictGraph<'a, const D: usize> {
    g: *mut ffi::gs_graph,
    _k: &'a Kriging<D>,
}

impl<const D: usize> PredictGraph<'_, D> {
    /// Enqueues the captured call on `stream` (null: the null stream).
    ///
    /// # Safety
    /// `stream` must be a stream of the predictor's GPU; launches of one
    /// graph must not overlap.
    pub unsafe fn launch(&self, stream: *mut c_void) -> Result<()> {
        status_to_result(ffi::gs_graph_launch(self.g, stream), [0, 0])
    }
}

impl<const D: usize> Drop for PredictGraph<'_, D> {
    fn drop(&mut self) {
        unsafe { ffi::gs_graph_free(self.g) }
    }
}

impl Kriging<2> {
    /// Kriging::predict_grid (kriging.rs:783-786): (values, variances) in
    /// storage order.
    pub fn predict_grid(&self, grid: &Grid2D) -> (Vec<f64>, Vec<f64>) {
        self.predict_many(&grid.centers()).into_iter().map(|e| (e.value, e.variance)).unzip()
    }

    /// Kriging::predict_block_grid (kriging.rs:790-803).
    pub fn predict_block_grid(&self, grid: &Grid2D, block_size: [f64; 2], discr: [usize; 2])
                              -> Result<(Vec<f64>, Vec<f64>)> {
        let offsets = block_offsets(block_size, discr)?;
        Ok(self
            .predict_block_many(&grid.centers(), &offsets)
            .into_iter()
            .map(|e| (e.value, e.variance))
            .unzip())
    }
}

impl<const D: usize> Drop for Kriging<D> {
    fn drop(&mut self) {
        unsafe { ffi::gs_kriging_free(self.h) }
    }
}

fn cv<const D: usize>(data: &PointSet<D>, predicted: Vec<f64>, variance: Vec<f64>) -> CvResult {
    CvResult { observed: data.values().to_vec(), predicted, variance }
}

/// leave_one_out (validation.rs:136-159).
pub fn leave_one_out<const D: usize>(ctx: &Context, data: &PointSet<D>, model: &VariogramModel,
                                     config: &KrigingConfig) -> Result<CvResult> {
    let (m, c) = (to_gs_model(model)?, to_gs_config(config));
    let n = data.len();
    let (mut p, mut v) = (vec![0.0; n], vec![0.0; n]);
    let st = unsafe {
        ffi::gs_cv_leave_one_out(ctx.h, D as i32, flat(data.coords()).as_ptr(),
                                 data.values().as_ptr(), n as i64, &m, &c, p.as_mut_ptr(),
                                 v.as_mut_ptr())
    };
    status_to_result(st, [0, 0])?;
    Ok(cv(data, p, v))
}

/// leave_one_out_with_drift (validation.rs:163-194).
pub fn leave_one_out_with_drift<const D: usize>(ctx: &Context, data: &PointSet<D>,
                                                drift_data: &[Vec<f64>], model: &VariogramModel,
                                                config: &KrigingConfig) -> Result<CvResult> {
    let width = drift_data.first().map(|r| r.len()).unwrap_or(0);
    let d = rows(drift_data, width)?;
    let (m, c) = (to_gs_model(model)?, to_gs_config(config));
    let n = data.len();
    let (mut p, mut v) = (vec![0.0; n], vec![0.0; n]);
    let st = unsafe {
        ffi::gs_cv_leave_one_out_with_drift(ctx.h, D as i32, flat(data.coords()).as_ptr(),
                                            data.values().as_ptr(), n as i64, &m, &c,
                                            width as i32, d.as_ptr(), p.as_mut_ptr(),
                                            v.as_mut_ptr())
    };
    status_to_result(st, [0, 0])?;
    Ok(cv(data, p, v))
}

/// k_fold (validation.rs:272-292).
pub fn k_fold<const D: usize>(ctx: &Context, data: &PointSet<D>, model: &VariogramModel,
                              config: &KrigingConfig, k: usize, seed: u64) -> Result<CvResult> {
    let (m, c) = (to_gs_model(model)?, to_gs_config(config));
    let n = data.len();
    let (mut p, mut v) = (vec![0.0; n], vec![0.0; n]);
    let st = unsafe {
        ffi::gs_cv_k_fold(ctx.h, D as i32, flat(data.coords()).as_ptr(), data.values().as_ptr(),
                          n as i64, &m, &c, k as i64, seed, p.as_mut_ptr(), v.as_mut_ptr())
    };
    status_to_result(st, [0, 0])?;
    Ok(cv(data, p, v))
}

/// block_cv (validation.rs:348-373).
pub fn block_cv<const D: usize>(ctx: &Context, data: &PointSet<D>, model: &VariogramModel,
                                config: &KrigingConfig, blocks_per_dim: [usize; D])
                                -> Result<CvResult> {
    let (m, c) = (to_gs_model(model)?, to_gs_config(config));
    let n = data.len();
    let b: Vec<i64> = blocks_per_dim.iter().map(|&x| x as i64).collect();
    let (mut p, mut v) = (vec![0.0; n], vec![0.0; n]);
    let st = unsafe {
        ffi::gs_cv_block(ctx.h, D as i32, flat(data.coords()).as_ptr(), data.values().as_ptr(),
                         n as i64, &m, &c, b.as_ptr(), p.as_mut_ptr(), v.as_mut_ptr())
    };
    status_to_result(st, [0, 0])?;
    Ok(cv(data, p, v))
}

/// VecchiaPlan (vecchia.rs:223-228) with coordinates, ordering and
/// conditioning sets resident on the GPU.
pub struct VecchiaPlan<const D: usize> {
    h: *mut ffi::gs_vplan,
    n: usize,
    m: usize,
}

unsafe impl<const D: usize> Send for VecchiaPlan<D> {}
unsafe impl<const D: usize> Sync for VecchiaPlan<D> {}

/// maxmin_order (vecchia.rs:144-215), computed on the context's GPU.
pub fn maxmin_order<const D: usize>(ctx: &Context, coords: &[[f64; D]]) -> Result<Vec<usize>> {
    let mut o = vec![0i64; coords.len()];
    let st = unsafe {
        ffi::gs_maxmin_order_gpu(ctx.h, D as i32, flat(coords).as_ptr(), coords.len() as i64,
                                 o.as_mut_ptr(), ptr::null_mut())
    };
    status_to_result(st, [0, 0])?;
    Ok(o.into_iter().map(|x| x as usize).collect())
}

/// vecchia_plan (vecchia.rs:232-282): maxmin ordering (order = None) or the
/// given permutation, then each point's m nearest predecessors.
pub fn vecchia_plan<const D: usize>(ctx: &Context, coords: &[[f64; D]], m: usize,
                                    order: Option<&[usize]>) -> Result<VecchiaPlan<D>> {
    let o: Option<Vec<i64>> = order.map(|o| o.iter().map(|&x| x as i64).collect());
    let mut h = ptr::null_mut();
    let st = unsafe {
        ffi::gs_vecchia_plan_new(ctx.h, D as i32, flat(coords).as_ptr(), coords.len() as i64,
                                 m as i64, o.as_ref().map_or(ptr::null(), |v| v.as_ptr()), &mut h)
    };
    status_to_result(st, [0, 0])?;
    Ok(VecchiaPlan { h, n: coords.len(), m })
}

impl<const D: usize> VecchiaPlan<D> {
    /// The ordering and each ordered point's conditioning indices.
    pub fn order_and_neighbours(&self) -> Result<(Vec<usize>, Vec<Vec<usize>>)> {
        let mut order = vec![0i64; self.n];
        let mut nb = vec![0i32; self.n * self.m];
        let mut cnt = vec![0i32; self.n];
        let st = unsafe {
            ffi::gs_vecchia_plan_get(self.h, order.as_mut_ptr(), nb.as_mut_ptr(), cnt.as_mut_ptr())
        };
        status_to_result(st, [0, 0])?;
        let sets = (0..self.n)
            .map(|k| nb[k * self.m..k * self.m + cnt[k] as usize].iter().map(|&i| i as usize).collect())
            .collect();
        Ok((order.into_iter().map(|x| x as usize).collect(), sets))
    }

    /// loglik_with_plan (vecchia.rs:285-336) for many models at once: the
    /// batched extension the reference's MLE loop needs.
    pub fn loglik_batch(&self, values: &[f64], models: &[VariogramModel], precision: Precision)
                        -> Result<Vec<Result<f64>>> {
        let gm: Vec<ffi::GsModel> = models.iter().map(to_gs_model).collect::<Result<_>>()?;
        let mut out = vec![0.0; models.len()];
        let mut st = vec![0i32; models.len()];
        let rc = unsafe {
            ffi::gs_vecchia_loglik_batch(self.h, values.as_ptr(), gm.as_ptr(), gm.len() as i32,
                                         precision.code(), out.as_mut_ptr(), st.as_mut_ptr())
        };
        status_to_result(rc, [0, 0])?;
        Ok(out.into_iter().zip(st).map(|(v, s)| status_to_result(s, [0, 0]).map(|_| v)).collect())
    }

    /// loglik_with_plan (vecchia.rs:285-336) for one model.
    pub fn loglik(&self, values: &[f64], model: &VariogramModel) -> Result<f64> {
        self.loglik_batch(values, std::slice::from_ref(model), Precision::Fp64)?.remove(0)
    }

    /// loglik_grad_with_plan (vecchia.rs:416-566): (loglik, d/dnugget,
    /// d/dpsill, d/drange) per model.
    pub fn loglik_grad_batch(&self, values: &[f64], models: &[VariogramModel])
                             -> Result<Vec<Result<(f64, [f64; 3])>>> {
        let gm: Vec<ffi::GsModel> = models.iter().map(to_gs_model).collect::<Result<_>>()?;
        let mut ll = vec![0.0; models.len()];
        let mut grad = vec![0.0; 3 * models.len()];
        let mut st = vec![0i32; models.len()];
        let rc = unsafe {
            ffi::gs_vecchia_loglik_grad_batch(self.h, values.as_ptr(), gm.as_ptr(),
                                              gm.len() as i32, ffi::GS_PREC_FP64,
                                              ll.as_mut_ptr(), grad.as_mut_ptr(), st.as_mut_ptr())
        };
        status_to_result(rc, [0, 0])?;
        Ok((0..models.len())
            .map(|i| status_to_result(st[i], [0, 0])
                .map(|_| (ll[i], [grad[3 * i], grad[3 * i + 1], grad[3 * i + 2]])))
            .collect())
    }

    /// mle_fit_with_plan (vecchia.rs:952-1081) of one structure of `kind`
    /// plus a nugget.
    pub fn mle(&self, values: &[f64], kind: ModelKind) -> Result<VecchiaFit> {
        let (code, param) = kind_code(kind);
        let mut fit = ffi::GsModel::default();
        let mut ll = 0.0;
        let st = unsafe { ffi::gs_vecchia_mle(self.h, values.as_ptr(), code, param, &mut fit, &mut ll) };
        status_to_result(st, [0, 0])?;
        Ok(VecchiaFit { model: from_gs_model(&fit)?, loglik: ll })
    }
}

impl<const D: usize> Drop for VecchiaPlan<D> {
    fn drop(&mut self) {
        unsafe { ffi::gs_vecchia_plan_free(self.h) }
    }
}

/// vecchia_loglik (vecchia.rs:874-884).
pub fn vecchia_loglik<const D: usize>(ctx: &Context, data: &PointSet<D>, model: &VariogramModel,
                                      m: usize, order: Option<&[usize]>) -> Result<f64> {
    let gm = to_gs_model(model)?;
    let o: Option<Vec<i64>> = order.map(|o| o.iter().map(|&x| x as i64).collect());
    let mut out = 0.0;
    let st = unsafe {
        ffi::gs_vecchia_loglik(ctx.h, D as i32, flat(data.coords()).as_ptr(),
                               data.values().as_ptr(), data.len() as i64, &gm, m as i64,
                               o.as_ref().map_or(ptr::null(), |v| v.as_ptr()), &mut out)
    };
    status_to_result(st, [0, 0])?;
    Ok(out)
}

/// vecchia_mle (vecchia.rs:921-932).
pub fn vecchia_mle<const D: usize>(ctx: &Context, data: &PointSet<D>, kind: ModelKind, m: usize,
                                   order: Option<&[usize]>) -> Result<VecchiaFit> {
    vecchia_plan(ctx, data.coords(), m, order)?.mle(data.values(), kind)
}
