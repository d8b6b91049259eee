/*
 * refcpu — CPU oracle for the geostat-hip kriging / Vecchia hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (franciscoparrao/geostat-rs, crates/geostat-core) that the HIP
 * implementation's results are compared with.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  The product library
 * (libgeostat_hip.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Compiled with
 * -O2 -ffp-contract=off so that floating-point expressions are evaluated in
 * the same order and with the same roundings as the Rust source (rustc never
 * contracts a*b+c into an FMA).
 *
 * Parity pins (see tests/test_oracle.py): reference inline test constants
 * (gstat / R values), brute-force kNN and maxmin, dense-MVN loglik, closed
 * forms.  Functions with no such pin are marked "parity unpinned" below.
 */
#ifndef GEOSTAT_REFCPU_H
#define GEOSTAT_REFCPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* model kinds: same numbering as include/geostat_hip.h (gs_kind) */
enum { RC_SPHERICAL = 0, RC_EXPONENTIAL, RC_GAUSSIAN, RC_MATERN15, RC_MATERN25,
       RC_MATERN, RC_CIRCULAR, RC_STABLE, RC_HOLE, RC_WAVE, RC_POWER };

enum { RC_OK = 0, RC_INSUFFICIENT_DATA = 1, RC_DIMENSION_MISMATCH = 2,
       RC_SINGULAR_SYSTEM = 3, RC_INVALID_PARAMETER = 4, RC_NO_NEIGHBORS = 5,
       RC_DUPLICATE_POINTS = 6 };

enum { RC_SIMPLE = 0, RC_ORDINARY = 1, RC_UNIVERSAL = 2, RC_EXTERNAL_DRIFT = 3 };

typedef struct {
    double azimuth_deg, ratio, ratio_z, dip_deg, rake_deg;
} rc_aniso;

typedef struct {
    int32_t kind;
    double param;      /* nu (Matern), alpha (Stable), theta (Power) */
    double sill;
    double range;
    int32_t has_anis;
    rc_aniso anis;
} rc_structure;

#define RC_MAX_STRUCT 8
typedef struct {
    double nugget;
    int32_t n_structures;
    rc_structure s[RC_MAX_STRUCT];
} rc_model;

typedef struct {
    int32_t method;          /* RC_SIMPLE / RC_ORDINARY / RC_UNIVERSAL */
    double simple_mean;
    int32_t degree;          /* universal drift degree; n_vars for external drift */
    int64_t max_neighbors;   /* -1 = None */
    int64_t min_neighbors;   /* -1 = None */
    int64_t max_per_octant;  /* -1 = None */
    int32_t has_radius;
    double radius;
    int32_t has_anis_search;
    rc_aniso anis_search;
} rc_kconfig;

/* ---- rng.rs ---- */
typedef struct { uint64_t s[4]; int32_t has_spare; double spare; } rc_rng;
void rc_rng_new(rc_rng *r, uint64_t seed);
uint64_t rc_rng_next_u64(rc_rng *r);
double rc_rng_uniform(rc_rng *r);
double rc_rng_normal(rc_rng *r);

/* ---- variogram/model.rs + bessel.rs ---- */
double rc_bessel_k(double nu, double x);
double rc_gamma_fn(double x);
double rc_shape_g(int32_t kind, double param, double h, double a);
void rc_transform_vector(const rc_aniso *a, int32_t D, const double *in, double *out);
void rc_rotation_matrix_3d(const rc_aniso *a, double rot[9]);
double rc_gamma_dh(const rc_model *m, int32_t D, const double *dh);
double rc_total_sill(const rc_model *m);
int32_t rc_model_validate(const rc_model *m, int32_t D);

/* ---- linalg.rs ---- */
int32_t rc_lu_solve(double *a, int32_t n, double *b);
int32_t rc_cholesky_solve(double *a, int32_t n, double *b);

/* ---- search.rs ---- */
int64_t rc_kdtree_knn(int32_t D, const double *coords, int64_t n,
                      const double *targets, int64_t T, int64_t k,
                      int32_t has_radius, double radius,
                      int32_t *out_idx, int32_t *out_count);
int32_t rc_kdtree_layout(int32_t D, const double *coords, int64_t n, uint32_t *idx_out);

/* ---- kriging.rs ---- */
int32_t rc_kriging_predict_many(int32_t D, const double *coords, const double *values, int64_t n,
                                const rc_model *model, const rc_kconfig *cfg,
                                const double *targets, int64_t T, int32_t nthreads,
                                double *value, double *variance, double *lagrange,
                                int32_t *status, int64_t *dup_pair);
/* measurement error / external drift / block kriging (kriging.rs:181-260,
 * 494-522, 656-672, 678-777): errors (n) or NULL; drift_data n x n_vars or
 * NULL; tdrift T x n_vars (predict_with_drift) or NULL; block != 0 runs
 * predict_block with the n_off offsets at every target. */
int32_t rc_kriging_predict_many_ex(int32_t D, const double *coords, const double *values,
                                   int64_t n, const rc_model *model, const rc_kconfig *cfg,
                                   const double *errors, int32_t n_vars, const double *drift_data,
                                   const double *targets, int64_t T, const double *tdrift,
                                   const double *offsets, int64_t n_off, int32_t block,
                                   int32_t nthreads, double *value, double *variance,
                                   double *lagrange, int32_t *status, int64_t *dup_pair);
int32_t rc_kriging_neighbors(int32_t D, const double *coords, const double *values, int64_t n,
                             const rc_model *model, const rc_kconfig *cfg,
                             const double *targets, int64_t T, int64_t max_out,
                             int32_t *idx_out, int32_t *count_out, int32_t nthreads);

/* ---- vecchia.rs ---- */
int32_t rc_maxmin_order(int32_t D, const double *coords, int64_t n, int64_t *order);
int32_t rc_maxmin_order_brute(int32_t D, const double *coords, int64_t n, int64_t *order);
int32_t rc_vecchia_plan(int32_t D, const double *coords, int64_t n, int64_t m,
                        const int64_t *order_in, int64_t *order_out,
                        int32_t *nb /* n*m, row k = ordered point k */, int32_t *nb_count);
int32_t rc_vecchia_loglik_plan(int32_t D, const double *coords, const double *values, int64_t n,
                               const rc_model *model, const int64_t *order,
                               const int32_t *nb, const int32_t *nb_count, int64_t m,
                               double *out);
int32_t rc_vecchia_loglik_grad_plan(int32_t D, const double *coords, const double *values,
                                    int64_t n, int32_t kind, double nugget, double psill,
                                    double range, const int64_t *order, const int32_t *nb,
                                    const int32_t *nb_count, int64_t m, double *ll_out,
                                    double *grad_out);
/* mle_fit_with_plan (vecchia.rs:952-1081), group_size 1: fit = (nugget,
 * psill, range); evals (may be NULL) counts objective + gradient calls. */
int32_t rc_vecchia_mle_plan(int32_t D, const double *coords, const double *values, int64_t n,
                            int32_t kind, double param, const int64_t *order, const int32_t *nb,
                            const int32_t *nb_count, int64_t m, double *fit, double *loglik,
                            int64_t *evals);
int32_t rc_dense_mvn_loglik(int32_t D, const double *coords, const double *values, int64_t n,
                            const rc_model *model, double *out);

/* ---- data.rs ---- */
int32_t rc_duplicate_pair(int32_t D, const double *coords, int64_t n, int64_t *pair);

#ifdef __cplusplus
}
#endif
#endif
