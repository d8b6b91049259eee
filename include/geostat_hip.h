/*
 * geostat_hip.h — C ABI of libgeostat_hip.so, the MI355X (gfx950) moving-
 * neighbourhood kriging predictor and Vecchia log-likelihood.
 *
 * This is the drop-in boundary for geostat-core's rayon `parallel` path:
 * every entry point below names the reference interface it replaces
 * (franciscoparrao/geostat-rs, paths under crates/geostat-core/src/).
 *
 * Conventions
 *  - The caller owns every host buffer; handles own their device memory
 *    until *_free.  No C++ exception crosses this boundary.
 *  - Return values are gs_status codes.  Codes 1..6 mirror GeostatError
 *    (error.rs:15-40); gs_last_error() returns a thread-local message.
 *  - Per-target failures are NaN estimates plus a per-target status code,
 *    never a call failure (kriging.rs:130-134, 648-651).
 *  - "None" in an Option<usize> config field is GS_NONE (-1).
 *  - Unsupported configurations return GS_ERR_INVALID_PARAMETER with a
 *    message starting "unsupported on GPU"; they never diverge silently.
 */
#ifndef GEOSTAT_HIP_H
#define GEOSTAT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GS_ABI_VERSION 2
#define GS_MAX_STRUCTURES 8
#define GS_NONE (-1)

typedef enum {
    GS_OK = 0,
    GS_ERR_INSUFFICIENT_DATA = 1,   /* GeostatError::InsufficientData */
    GS_ERR_DIMENSION_MISMATCH = 2,  /* GeostatError::DimensionMismatch */
    GS_ERR_SINGULAR_SYSTEM = 3,     /* GeostatError::SingularSystem */
    GS_ERR_INVALID_PARAMETER = 4,   /* GeostatError::InvalidParameter */
    GS_ERR_NO_NEIGHBORS = 5,        /* GeostatError::NoNeighbors */
    GS_ERR_DUPLICATE_POINTS = 6,    /* GeostatError::DuplicatePoints */
    GS_ERR_DEVICE = 100,            /* HIP runtime failure */
    GS_ERR_OUT_OF_MEMORY = 101      /* device allocation failed */
} gs_status;

/* ModelKind (variogram/model.rs:47-95).  `param` carries nu / alpha / theta. */
typedef enum {
    GS_SPHERICAL = 0, GS_EXPONENTIAL = 1, GS_GAUSSIAN = 2, GS_MATERN15 = 3,
    GS_MATERN25 = 4, GS_MATERN = 5, GS_CIRCULAR = 6, GS_STABLE = 7,
    GS_HOLE = 8, GS_WAVE = 9, GS_POWER = 10
} gs_kind;

/* KrigingMethod (kriging.rs:25-44); external drift is not exposed. */
/* KrigingMethod (kriging.rs:25-44).  GS_EXTERNAL_DRIFT handles are built by
 * gs_kriging_new_with_external_drift, which takes the covariate count. */
typedef enum { GS_SIMPLE = 0, GS_ORDINARY = 1, GS_UNIVERSAL = 2, GS_EXTERNAL_DRIFT = 3 } gs_method;

/* Arithmetic of the covariance build and solve. */
typedef enum {
    GS_PREC_FP64 = 0,          /* FP64 throughout */
    GS_PREC_FP32_REFINED = 1   /* FP32 factorization + FP64-residual refinement */
} gs_precision;

/* Anisotropy (variogram/model.rs:321-341). */
typedef struct {
    double azimuth_deg, ratio, ratio_z, dip_deg, rake_deg;
} gs_anisotropy;

/* Structure (variogram/model.rs:438-449). */
typedef struct {
    int32_t kind;        /* gs_kind */
    double param;        /* Matern nu, Stable alpha, Power theta; else ignored */
    double sill;
    double range;
    int32_t has_anis;
    gs_anisotropy anis;
} gs_structure;

/* VariogramModel (variogram/model.rs:543-549). */
typedef struct {
    double nugget;
    int32_t n_structures;
    gs_structure structures[GS_MAX_STRUCTURES];
} gs_model;

/* KrigingConfig (kriging.rs:73-115). */
typedef struct {
    int32_t method;            /* gs_method */
    double simple_mean;        /* KrigingMethod::Simple { mean } */
    int32_t degree;            /* KrigingMethod::Universal { degree } */
    int64_t max_neighbors;     /* Option<usize>: GS_NONE or >= 0 */
    int64_t min_neighbors;     /* Option<usize> */
    int64_t max_per_octant;    /* Option<usize> */
    int32_t has_search_radius;
    double search_radius;
    int32_t has_anisotropic_search;
    gs_anisotropy anisotropic_search;
} gs_kconfig;

typedef struct gs_ctx gs_ctx;
typedef struct gs_kriging gs_kriging;
typedef struct gs_vplan gs_vplan;

/* ---- context ------------------------------------------------------ */

/* Opens `device` and creates the context's default stream.  (No reference
 * counterpart: the rayon path has no device state.) */
int32_t gs_ctx_create(int32_t device, gs_ctx **out);
void gs_ctx_destroy(gs_ctx *ctx);
int32_t gs_ctx_synchronize(gs_ctx *ctx);
/* Thread-local message of the last failing call on this thread. */
const char *gs_last_error(void);
/* Kernel family of the last solve this thread launched ("krige_mfma_kernel",
 * "krige_fp64_kernel", "krige_fp32r_kernel", "vecchia_mfma_kernel",
 * "vecchia_kernel", ...; "" before any).  No reference counterpart: a
 * measurement hook naming the kernel a roofline figure belongs to. */
const char *gs_last_solve_kernel(void);
int32_t gs_abi_version(void);
/* VariogramModel::new's checks (model.rs:554-618): nugget, sills, ranges,
 * kind parameters, anisotropy, positive total sill.  GS_OK or
 * GS_ERR_INVALID_PARAMETER (more than 8 structures is rejected too). */
int32_t gs_model_validate(const gs_model *model);

/* ---- kriging ------------------------------------------------------ */

/* Kriging::<D, VariogramModel>::new (kriging.rs:166-173 → build 262-408):
 * validates data, model and config, rejects duplicate locations
 * (data.rs:150-163; the first pair goes to dup_pair), builds the search
 * index on the (optionally anisotropic) search frame and keeps data and
 * index resident in HBM.  coords: n*dim AoS doubles; values: n doubles. */
int32_t gs_kriging_new(gs_ctx *ctx, int32_t dim, const double *coords, const double *values,
                       int64_t n, const gs_model *model, const gs_kconfig *config,
                       gs_kriging **out, int64_t dup_pair[2]);
void gs_kriging_free(gs_kriging *k);

/* Kriging::with_measurement_error (kriging.rs:181-218): errors[i] (n
 * doubles, finite, >= 0) is the error variance of datum i, added to the
 * data-data diagonal only.  External drift and Power models are rejected
 * as in the reference. */
int32_t gs_kriging_new_with_measurement_error(gs_ctx *ctx, int32_t dim, const double *coords,
                                              const double *values, int64_t n,
                                              const gs_model *model, const gs_kconfig *config,
                                              const double *errors, gs_kriging **out,
                                              int64_t dup_pair[2]);

/* Kriging::with_external_drift (kriging.rs:223-260): config->method must be
 * GS_EXTERNAL_DRIFT; drift_data is n*n_vars doubles (row i = covariates of
 * datum i), normalized per covariate as the reference does (349-364).  On
 * the GPU n_vars is 1..3. */
int32_t gs_kriging_new_with_external_drift(gs_ctx *ctx, int32_t dim, const double *coords,
                                           const double *values, int64_t n,
                                           const gs_model *model, const gs_kconfig *config,
                                           int32_t n_vars, const double *drift_data,
                                           gs_kriging **out, int64_t dup_pair[2]);

/* Kriging::predict_many (kriging.rs:648-652) over host buffers: targets is
 * T*dim AoS doubles; outputs have T entries each.  lagrange is NaN where the
 * reference returns None; status holds the per-target gs_status (0 = ok;
 * NaN estimate otherwise).  lagrange and status may be NULL. */
int32_t gs_kriging_predict_many(gs_kriging *k, const double *targets, int64_t T,
                                int32_t precision, double *value, double *variance,
                                double *lagrange, int32_t *status);

/* Kriging::predict_many_with_drift (kriging.rs:656-672): drifts is T*n_vars
 * doubles (the covariates at each target).  On a handle without external
 * drift every target fails with GS_ERR_INVALID_PARAMETER, as the
 * reference's predict_with_drift does per target (507-512); predict_many
 * on an external-drift handle likewise fails every target (494-501). */
int32_t gs_kriging_predict_many_with_drift(gs_kriging *k, const double *targets, int64_t T,
                                           const double *drifts, int32_t precision,
                                           double *value, double *variance, double *lagrange,
                                           int32_t *status);

/* Kriging::predict_block (kriging.rs:678-777) at T block centres (T*dim
 * AoS), each block discretized at centre + offsets[u] (n_off*dim AoS; the
 * reference's block_offsets, 817-838, gives the regular grid used by
 * predict_block_grid, 790-803).  Simple and ordinary kriging only: other
 * methods, and n_off = 0, fail every target with GS_ERR_INVALID_PARAMETER
 * (the per-target Err that predict_block_grid maps to NaN). */
int32_t gs_kriging_predict_block_many(gs_kriging *k, const double *centers, int64_t T,
                                      const double *offsets, int32_t n_off, int32_t precision,
                                      double *value, double *variance, double *lagrange,
                                      int32_t *status);

/* Same as gs_kriging_predict_many with every buffer already in device
 * memory, enqueued on `stream` (a hipStream_t; NULL is the null stream, as
 * in HIP's own calls, so work the caller queued there, such as filling
 * the buffers, comes first).  Asynchronous: returns after enqueueing.
 * Workspace comes from the handle's cache, so calls on one handle must not
 * overlap. */
int32_t gs_kriging_predict_many_device(gs_kriging *k, const double *d_targets, int64_t T,
                                       int32_t precision, double *d_value, double *d_variance,
                                       double *d_lagrange, int32_t *d_status, void *stream);

/* gs_kriging_predict_many_device without any device allocation (no
 * reference counterpart: for callers that cannot allocate, such as code under
 * HIP graph capture, tight serving loops or harnesses that forbid hipMalloc).
 * The workspace comes from a fixed 64 MiB scratch per GPU, allocated when the
 * library is loaded and held until the process ends; a call whose workspace
 * would exceed it (about 440,000 targets at 32 neighbours) fails with
 * GS_ERR_INVALID_PARAMETER before any work is enqueued, and so does a handle
 * with a global neighbourhood or neighbourhoods above 64 points.  Every call
 * on one GPU shares that GPU's
 * scratch: calls that overlap on one GPU race (serialize them).  The kd-tree
 * replay of flagged targets stays on `stream`.  Results are bit-identical to
 * gs_kriging_predict_many_device's. */
int32_t gs_kriging_predict_many_noalloc(gs_kriging *k, const double *d_targets, int64_t T,
                                        int32_t precision, double *d_value, double *d_variance,
                                        double *d_lagrange, int32_t *d_status, void *stream);

/* A predict call captured as a HIP graph (no reference counterpart: the
 * rayon par_map of predict_many, parallel.rs:9-17, replayed as one graph
 * launch for callers that predict the same number of targets over and over,
 * such as a serving loop or a simulation refreshing a field).
 * gs_kriging_graph_create runs the call of gs_kriging_predict_many_device
 * once (its results land in the output buffers) and records its launches
 * over these buffers and a workspace owned by the graph; each
 * gs_graph_launch then enqueues the whole call on `stream` as one graph
 * launch, reading whatever the target buffer holds at that point.  Results
 * are bit-identical to gs_kriging_predict_many_device's on the same targets.
 * The predictor and the buffers must outlive the graph; launches of one graph
 * must not overlap (they share its workspace); the predictor's own calls may
 * run beside them.  Neighbourhoods above 64 points are refused
 * (GS_ERR_INVALID_PARAMETER), since their work is sized on the host. */
typedef struct gs_graph gs_graph;
int32_t gs_kriging_graph_create(gs_kriging *k, const double *d_targets, int64_t T,
                                int32_t precision, double *d_value, double *d_variance,
                                double *d_lagrange, int32_t *d_status, gs_graph **out);
int32_t gs_graph_launch(gs_graph *g, void *stream);
void gs_graph_free(gs_graph *g);

/* Kriging::neighbors (kriging.rs:441-490) for T targets: idx is T*max_out
 * int32 (original data indices in the reference's order, -1 padded),
 * count has T entries (the neighbourhood size before any min_neighbors
 * check).  max_out must be >= the configured neighbourhood bound. */
int32_t gs_kriging_neighbors(gs_kriging *k, const double *targets, int64_t T, int64_t max_out,
                             int32_t *idx, int32_t *count);

/* Largest neighbourhood one target can receive under this handle's config. */
int64_t gs_kriging_max_neighbors(const gs_kriging *k);

/* Per-phase device timing (no reference counterpart; a measurement hook).
 * While enabled, every predict call on the handle records HIP events on its
 * launch stream around the neighbour search and around the solves.
 * gs_kriging_timing_summary waits for the recorded events, returns the
 * number of calls and the summed milliseconds of each phase since the last
 * summary, and clears the record. */
int32_t gs_kriging_set_timing(gs_kriging *k, int32_t enable);
int32_t gs_kriging_timing_summary(gs_kriging *k, int64_t *calls, double *ms_search,
                                  double *ms_solve);

/* ---- cross-validation (validation.rs) ------------------------------- */
/* Each fold builds a predictor on the points outside it (the same checks
 * as gs_kriging_new) and predicts its members; predicted / variance are n
 * doubles in dataset order.  The first failing fold (and member) decides
 * the returned error, as the reference's sequential collect does.  At
 * least 3 points. */

/* leave_one_out (validation.rs:136-159). */
int32_t gs_cv_leave_one_out(gs_ctx *ctx, int32_t dim, const double *coords, const double *values,
                            int64_t n, const gs_model *model, const gs_kconfig *config,
                            double *predicted, double *variance);
/* leave_one_out_with_drift (validation.rs:163-194): drift_data n*n_vars. */
int32_t gs_cv_leave_one_out_with_drift(gs_ctx *ctx, int32_t dim, const double *coords,
                                       const double *values, int64_t n, const gs_model *model,
                                       const gs_kconfig *config, int32_t n_vars,
                                       const double *drift_data, double *predicted,
                                       double *variance);
/* k_fold (validation.rs:272-292): folds from the seeded shuffle of
 * fold_assignment (199-212), 2 <= k <= n. */
int32_t gs_cv_k_fold(gs_ctx *ctx, int32_t dim, const double *coords, const double *values,
                     int64_t n, const gs_model *model, const gs_kconfig *config, int64_t k,
                     uint64_t seed, double *predicted, double *variance);
/* block_cv (validation.rs:348-373): blocks_per_dim[dim] regular blocks over
 * the bounding box (spatial_block_assignment, 294-323); empty blocks are
 * skipped. */
int32_t gs_cv_block(gs_ctx *ctx, int32_t dim, const double *coords, const double *values,
                    int64_t n, const gs_model *model, const gs_kconfig *config,
                    const int64_t *blocks_per_dim, double *predicted, double *variance);

/* ---- Vecchia ------------------------------------------------------ */

/* maxmin_order (vecchia.rs:144-215): the maxmin (farthest-point) ordering of
 * n finite points, written to order[n].  Host computation, no device work:
 * the first point is the one nearest to the centroid, then each next point
 * is the one farthest from those already ordered, ties to the larger
 * index. */
int32_t gs_maxmin_order(int32_t dim, const double *coords, int64_t n, int64_t *order);
/* The same order computed on the device of ctx (maxmin_order,
 * vecchia.rs:144-215): rounds that each append the longest run of the
 * largest-key candidates that cannot lower each other's keys; the host
 * finishes the traversal when rounds append few points (tie-heavy inputs such
 * as lattices).  info (3 entries, may be NULL): rounds run on the device,
 * rounds that raised the candidate threshold instead of appending, points
 * ordered on the device.  Used by gs_vecchia_plan_new when order is NULL. */
int32_t gs_maxmin_order_gpu(gs_ctx *ctx, int32_t dim, const double *coords, int64_t n,
                            int64_t *order, int64_t *info);

/* vecchia_plan (vecchia.rs:232-282): maxmin ordering (order == NULL) or the
 * given permutation, then each ordered point's m nearest predecessors. */
int32_t gs_vecchia_plan_new(gs_ctx *ctx, int32_t dim, const double *coords, int64_t n, int64_t m,
                            const int64_t *order, gs_vplan **out);
void gs_vecchia_plan_free(gs_vplan *p);
/* Copies the plan out: order (n), nb (n*m, row k = ordered point k, -1
 * padded, original indices), nb_count (n).  Any pointer may be NULL. */
int32_t gs_vecchia_plan_get(const gs_vplan *p, int64_t *order, int32_t *nb, int32_t *nb_count);

/* Batched loglik_with_plan (vecchia.rs:285-336) for n_models models over
 * one plan (the reference's plan-reusing loglik is private; this is the
 * batched extension its MLE loop needs).  out[i] is model i's
 * log-likelihood; status[i] is GS_ERR_SINGULAR_SYSTEM when any conditional
 * of model i fails (the reference's whole-call Err), GS_OK otherwise. */
int32_t gs_vecchia_loglik_batch(gs_vplan *p, const double *values, const gs_model *models,
                                int32_t n_models, int32_t precision, double *out,
                                int32_t *status);

/* Batched loglik_grad_with_plan (vecchia.rs:416-566): log-likelihood and its
 * gradient in (nugget, psill, range) for n_models single-structure
 * isotropic models of the kinds with a closed-form range derivative
 * (spherical, exponential, Gaussian, Matern 3/2, 5/2; g_and_dg,
 * vecchia.rs:353-385).  loglik[i], grad[3*i + {0,1,2}] = d/dnugget,
 * d/dpsill, d/drange.  status[i]: GS_OK; GS_ERR_INVALID_PARAMETER for a
 * model the reference answers with None (other kinds, several structures,
 * anisotropy); GS_ERR_SINGULAR_SYSTEM as in gs_vecchia_loglik_batch.
 * Outputs of a model with a non-zero status are NaN.  Plans whose
 * conditioning sets exceed 32 points are rejected. */
int32_t gs_vecchia_loglik_grad_batch(gs_vplan *p, const double *values, const gs_model *models,
                                     int32_t n_models, int32_t precision, double *loglik,
                                     double *grad, int32_t *status);

/* vecchia_loglik (vecchia.rs:874-884): plan + one likelihood. */
int32_t gs_vecchia_loglik(gs_ctx *ctx, int32_t dim, const double *coords, const double *values,
                          int64_t n, const gs_model *model, int64_t m, const int64_t *order,
                          double *out);

/* Vecchia maximum-likelihood fit of one structure of `kind` (param: the
 * Matern nu or Stable alpha, ignored otherwise) plus a nugget over a built
 * plan: mle_fit_with_plan (vecchia.rs:952-1081) as called by vecchia_mle
 * (vecchia.rs:921-932).  Same parametrization (sqrt nugget, ln psill,
 * ln range), starts, range regularization and optimizers (BFGS for the
 * g_and_dg kinds, Nelder-Mead otherwise; optim.rs:6-288); the starts run in
 * lockstep and each round of their evaluations is one batched launch.
 * fit receives the fitted model (nugget, one isotropic structure), loglik
 * its Vecchia log-likelihood.  GS_ERR_INVALID_PARAMETER for Power, for
 * Circular in 3-D, and (BFGS kinds) for plans whose conditioning sets
 * exceed 32 points; GS_ERR_SINGULAR_SYSTEM when every start stays on a
 * singular covariance. */
int32_t gs_vecchia_mle(gs_vplan *p, const double *values, int32_t kind, double param,
                       gs_model *fit, double *loglik);

/* ---- request loop ------------------------------------------------- */
/* Serving: kriging and Vecchia-likelihood requests over predictors and
 * plans that stay resident in HBM (no reference counterpart; the
 * reference's callers run one request at a time).  n_slots worker slots
 * (1..16), each with its own stream and workspace; requests are admitted
 * first come first served when a slot is free and their device bytes fit
 * in mem_budget beside the running ones, and run concurrently.  Results
 * are identical to the same request made alone (deterministic kernels).
 * The caller keeps a request's buffers alive until gs_server_wait returns
 * for its ticket; destroy drains the queue. */
typedef struct gs_server gs_server;
int32_t gs_server_create(gs_ctx *ctx, int32_t n_slots, int64_t mem_budget, gs_server **out);
void gs_server_destroy(gs_server *s);
/* A predict_many request on a handle (not external drift) at `precision`
 * (gs_precision). */
int32_t gs_server_submit_krige(gs_server *s, gs_kriging *k, const double *targets, int64_t T,
                               int32_t precision, double *value, double *variance,
                               double *lagrange, int32_t *status, int64_t *ticket);
/* A gs_vecchia_loglik_batch request on a plan at `precision`. */
int32_t gs_server_submit_vecchia(gs_server *s, gs_vplan *p, const double *values,
                                 const gs_model *models, int32_t n_models, int32_t precision,
                                 double *out, int32_t *status, int64_t *ticket);
/* Waits for a ticket; returns the request's status; queue_ms / service_ms
 * (may be NULL) receive its time in the queue and on a slot. */
int32_t gs_server_wait(gs_server *s, int64_t ticket, double *queue_ms, double *service_ms);

#ifdef __cplusplus
}
#endif
#endif
