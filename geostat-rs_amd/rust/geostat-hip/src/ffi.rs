//! Raw declarations of libgeostat_hip.so, one for one with
//! include/geostat_hip.h (ABI version 2).  Handles are opaque pointers; every
//! function returns a gs_status code (0 = ok) unless stated otherwise.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_void};

pub const GS_ABI_VERSION: i32 = 2;
pub const GS_MAX_STRUCTURES: usize = 8;
pub const GS_NONE: i64 = -1;

pub const GS_OK: i32 = 0;
pub const GS_ERR_INSUFFICIENT_DATA: i32 = 1;
pub const GS_ERR_DIMENSION_MISMATCH: i32 = 2;
pub const GS_ERR_SINGULAR_SYSTEM: i32 = 3;
pub const GS_ERR_INVALID_PARAMETER: i32 = 4;
pub const GS_ERR_NO_NEIGHBORS: i32 = 5;
pub const GS_ERR_DUPLICATE_POINTS: i32 = 6;
pub const GS_ERR_DEVICE: i32 = 100;
pub const GS_ERR_OUT_OF_MEMORY: i32 = 101;

pub const GS_PREC_FP64: i32 = 0;
pub const GS_PREC_FP32_REFINED: i32 = 1;

pub const GS_SIMPLE: i32 = 0;
pub const GS_ORDINARY: i32 = 1;
pub const GS_UNIVERSAL: i32 = 2;
pub const GS_EXTERNAL_DRIFT: i32 = 3;

pub type gs_ctx = c_void;
pub type gs_kriging = c_void;
pub type gs_vplan = c_void;
pub type gs_server = c_void;
pub type gs_graph = c_void;

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct GsAnisotropy {
    pub azimuth_deg: f64,
    pub ratio: f64,
    pub ratio_z: f64,
    pub dip_deg: f64,
    pub rake_deg: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct GsStructure {
    pub kind: i32,
    pub param: f64,
    pub sill: f64,
    pub range: f64,
    pub has_anis: i32,
    pub anis: GsAnisotropy,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct GsModel {
    pub nugget: f64,
    pub n_structures: i32,
    pub structures: [GsStructure; GS_MAX_STRUCTURES],
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct GsKConfig {
    pub method: i32,
    pub simple_mean: f64,
    pub degree: i32,
    pub max_neighbors: i64,
    pub min_neighbors: i64,
    pub max_per_octant: i64,
    pub has_search_radius: i32,
    pub search_radius: f64,
    pub has_anisotropic_search: i32,
    pub anisotropic_search: GsAnisotropy,
}

#[link(name = "geostat_hip")]
extern "C" {
    // context
    pub fn gs_ctx_create(device: i32, out: *mut *mut gs_ctx) -> i32;
    pub fn gs_ctx_destroy(ctx: *mut gs_ctx);
    pub fn gs_ctx_synchronize(ctx: *mut gs_ctx) -> i32;
    pub fn gs_last_error() -> *const c_char;
    pub fn gs_last_solve_kernel() -> *const c_char;
    pub fn gs_abi_version() -> i32;
    pub fn gs_model_validate(model: *const GsModel) -> i32;

    // kriging
    pub fn gs_kriging_new(ctx: *mut gs_ctx, dim: i32, coords: *const f64, values: *const f64,
                          n: i64, model: *const GsModel, config: *const GsKConfig,
                          out: *mut *mut gs_kriging, dup_pair: *mut i64) -> i32;
    pub fn gs_kriging_free(k: *mut gs_kriging);
    pub fn gs_kriging_new_with_measurement_error(ctx: *mut gs_ctx, dim: i32, coords: *const f64,
                                                 values: *const f64, n: i64,
                                                 model: *const GsModel, config: *const GsKConfig,
                                                 errors: *const f64, out: *mut *mut gs_kriging,
                                                 dup_pair: *mut i64) -> i32;
    pub fn gs_kriging_new_with_external_drift(ctx: *mut gs_ctx, dim: i32, coords: *const f64,
                                              values: *const f64, n: i64, model: *const GsModel,
                                              config: *const GsKConfig, n_vars: i32,
                                              drift_data: *const f64, out: *mut *mut gs_kriging,
                                              dup_pair: *mut i64) -> i32;
    pub fn gs_kriging_predict_many(k: *mut gs_kriging, targets: *const f64, t: i64,
                                   precision: i32, value: *mut f64, variance: *mut f64,
                                   lagrange: *mut f64, status: *mut i32) -> i32;
    pub fn gs_kriging_predict_many_with_drift(k: *mut gs_kriging, targets: *const f64, t: i64,
                                              drifts: *const f64, precision: i32,
                                              value: *mut f64, variance: *mut f64,
                                              lagrange: *mut f64, status: *mut i32) -> i32;
    pub fn gs_kriging_predict_block_many(k: *mut gs_kriging, centers: *const f64, t: i64,
                                         offsets: *const f64, n_off: i32, precision: i32,
                                         value: *mut f64, variance: *mut f64,
                                         lagrange: *mut f64, status: *mut i32) -> i32;
    pub fn gs_kriging_predict_many_device(k: *mut gs_kriging, d_targets: *const f64, t: i64,
                                          precision: i32, d_value: *mut f64,
                                          d_variance: *mut f64, d_lagrange: *mut f64,
                                          d_status: *mut i32, stream: *mut c_void) -> i32;
    pub fn gs_kriging_predict_many_noalloc(k: *mut gs_kriging, d_targets: *const f64, t: i64,
                                           precision: i32, d_value: *mut f64,
                                           d_variance: *mut f64, d_lagrange: *mut f64,
                                           d_status: *mut i32, stream: *mut c_void) -> i32;
    pub fn gs_kriging_graph_create(k: *mut gs_kriging, d_targets: *const f64, t: i64,
                                   precision: i32, d_value: *mut f64, d_variance: *mut f64,
                                   d_lagrange: *mut f64, d_status: *mut i32,
                                   out: *mut *mut gs_graph) -> i32;
    pub fn gs_graph_launch(g: *mut gs_graph, stream: *mut c_void) -> i32;
    pub fn gs_graph_free(g: *mut gs_graph);
    pub fn gs_kriging_neighbors(k: *mut gs_kriging, targets: *const f64, t: i64, max_out: i64,
                                idx: *mut i32, count: *mut i32) -> i32;
    pub fn gs_kriging_max_neighbors(k: *const gs_kriging) -> i64;
    pub fn gs_kriging_set_timing(k: *mut gs_kriging, enable: i32) -> i32;
    pub fn gs_kriging_timing_summary(k: *mut gs_kriging, calls: *mut i64, ms_search: *mut f64,
                                     ms_solve: *mut f64) -> i32;

    // cross-validation
    pub fn gs_cv_leave_one_out(ctx: *mut gs_ctx, dim: i32, coords: *const f64,
                               values: *const f64, n: i64, model: *const GsModel,
                               config: *const GsKConfig, predicted: *mut f64,
                               variance: *mut f64) -> i32;
    pub fn gs_cv_leave_one_out_with_drift(ctx: *mut gs_ctx, dim: i32, coords: *const f64,
                                          values: *const f64, n: i64, model: *const GsModel,
                                          config: *const GsKConfig, n_vars: i32,
                                          drift_data: *const f64, predicted: *mut f64,
                                          variance: *mut f64) -> i32;
    pub fn gs_cv_k_fold(ctx: *mut gs_ctx, dim: i32, coords: *const f64, values: *const f64,
                        n: i64, model: *const GsModel, config: *const GsKConfig, k: i64,
                        seed: u64, predicted: *mut f64, variance: *mut f64) -> i32;
    pub fn gs_cv_block(ctx: *mut gs_ctx, dim: i32, coords: *const f64, values: *const f64,
                       n: i64, model: *const GsModel, config: *const GsKConfig,
                       blocks_per_dim: *const i64, predicted: *mut f64,
                       variance: *mut f64) -> i32;

    // Vecchia
    pub fn gs_maxmin_order(dim: i32, coords: *const f64, n: i64, order: *mut i64) -> i32;
    pub fn gs_maxmin_order_gpu(ctx: *mut gs_ctx, dim: i32, coords: *const f64, n: i64,
                               order: *mut i64, info: *mut i64) -> i32;
    pub fn gs_vecchia_plan_new(ctx: *mut gs_ctx, dim: i32, coords: *const f64, n: i64, m: i64,
                               order: *const i64, out: *mut *mut gs_vplan) -> i32;
    pub fn gs_vecchia_plan_free(p: *mut gs_vplan);
    pub fn gs_vecchia_plan_get(p: *const gs_vplan, order: *mut i64, nb: *mut i32,
                               nb_count: *mut i32) -> i32;
    pub fn gs_vecchia_loglik_batch(p: *mut gs_vplan, values: *const f64, models: *const GsModel,
                                   n_models: i32, precision: i32, out: *mut f64,
                                   status: *mut i32) -> i32;
    pub fn gs_vecchia_loglik_grad_batch(p: *mut gs_vplan, values: *const f64,
                                        models: *const GsModel, n_models: i32, precision: i32,
                                        loglik: *mut f64, grad: *mut f64,
                                        status: *mut i32) -> i32;
    pub fn gs_vecchia_loglik(ctx: *mut gs_ctx, dim: i32, coords: *const f64, values: *const f64,
                             n: i64, model: *const GsModel, m: i64, order: *const i64,
                             out: *mut f64) -> i32;
    pub fn gs_vecchia_mle(p: *mut gs_vplan, values: *const f64, kind: i32, param: f64,
                          fit: *mut GsModel, loglik: *mut f64) -> i32;

    // request loop
    pub fn gs_server_create(ctx: *mut gs_ctx, n_slots: i32, mem_budget: i64,
                            out: *mut *mut gs_server) -> i32;
    pub fn gs_server_destroy(s: *mut gs_server);
    pub fn gs_server_submit_krige(s: *mut gs_server, k: *mut gs_kriging, targets: *const f64,
                                  t: i64, precision: i32, value: *mut f64, variance: *mut f64,
                                  lagrange: *mut f64, status: *mut i32,
                                  ticket: *mut i64) -> i32;
    pub fn gs_server_submit_vecchia(s: *mut gs_server, p: *mut gs_vplan, values: *const f64,
                                    models: *const GsModel, n_models: i32, precision: i32,
                                    out: *mut f64, status: *mut i32, ticket: *mut i64) -> i32;
    pub fn gs_server_wait(s: *mut gs_server, ticket: i64, queue_ms: *mut f64,
                          service_ms: *mut f64) -> i32;
}
