/*
 * refcpu.c — CPU oracle (TEST INFRASTRUCTURE ONLY; see refcpu.h).
 *
 * Plain-C restatement of geostat-core's kriging / Vecchia hot path.
 * All reference citations are paths under crates/geostat-core/src/ of
 * franciscoparrao/geostat-rs.  Build: oracle/Makefile (-O2 -ffp-contract=off).
 */
#define _GNU_SOURCE
#include "refcpu.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* rng.rs:8-75 — splitmix64 seeding, xoshiro256++, 53-bit uniform,     */
/* Box–Muller normal with cached spare (r·cos returned first).          */
/* ------------------------------------------------------------------ */
static uint64_t splitmix64(uint64_t *state) {            /* rng.rs:8-14 */
    *state += 0x9E3779B97F4A7C15ULL;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void rc_rng_new(rc_rng *r, uint64_t seed) {              /* rng.rs:25-37 */
    uint64_t sm = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&sm);
    r->has_spare = 0;
    r->spare = 0.0;
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t rc_rng_next_u64(rc_rng *r) {                    /* rng.rs:40-53 */
    uint64_t *s = r->s;
    uint64_t result = rotl(s[0] + s[3], 23) + s[0];
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

double rc_rng_uniform(rc_rng *r) {                       /* rng.rs:56-58 */
    return (double)(rc_rng_next_u64(r) >> 11) * (1.0 / 9007199254740992.0);
}

double rc_rng_normal(rc_rng *r) {                        /* rng.rs:61-75 */
    if (r->has_spare) {
        r->has_spare = 0;
        return r->spare;
    }
    for (;;) {
        double u1 = rc_rng_uniform(r);
        if (u1 > 0.0) {
            double u2 = rc_rng_uniform(r);
            double rr = sqrt(-2.0 * log(u1));
            double ang = 2.0 * M_PI * u2;
            double s = sin(ang), c = cos(ang);
            r->spare = rr * s;
            r->has_spare = 1;
            return rr * c;
        }
    }
}

/* ------------------------------------------------------------------ */
/* variogram/bessel.rs:29-123                                          */
/* ------------------------------------------------------------------ */
static void gauss_legendre(int n, double *x, double *w) { /* bessel.rs:29-57 */
    int m = (n + 1) / 2;
    for (int i = 0; i < m; i++) {
        double z = cos((M_PI * ((double)i + 0.75)) / ((double)n + 0.5));
        for (;;) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 0; j < n; j++) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * (double)j + 1.0) * z * p2 - (double)j * p3) / ((double)j + 1.0);
            }
            double pp = (double)n * (z * p1 - p2) / (z * z - 1.0);
            double z1 = z;
            z -= p1 / pp;
            if (fabs(z - z1) <= 1e-15) {
                w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
                x[i] = -z;
                w[n - 1 - i] = w[i];
                x[n - 1 - i] = z;
                break;
            }
        }
    }
}

static double GL60X[60], GL60W[60], GL40X[40], GL40W[40];
static pthread_once_t gl_once = PTHREAD_ONCE_INIT;
static void gl_init(void) {
    gauss_legendre(60, GL60X, GL60W);
    gauss_legendre(40, GL40X, GL40W);
}

/* Rust std f64::acosh: ln(x + sqrt(x-1)*sqrt(x+1)), NaN below 1. */
static double rust_acosh(double x) {
    if (x < 1.0) return NAN;
    return log(x + sqrt(x - 1.0) * sqrt(x + 1.0));
}

double rc_bessel_k(double nu, double x) {                /* bessel.rs:70-95 */
    pthread_once(&gl_once, gl_init);
    nu = fabs(nu);
    double integral = 0.0;
    for (int i = 0; i < 60; i++) {
        double t = 0.5 * (GL60X[i] + 1.0);
        integral += GL60W[i] * (exp(-x * cosh(t)) * cosh(nu * t));
    }
    integral *= 0.5;
    double xm = x > 1e-12 ? x : 1e-12;
    double tm = rust_acosh(60.0 / xm);
    double t_max = fmax(tm, 4.0) + 1.0;
    double integral2 = 0.0;
    for (int i = 0; i < 40; i++) {
        double t = 1.0 + 0.5 * (t_max - 1.0) * (GL40X[i] + 1.0);
        integral2 += GL40W[i] * (exp(-x * cosh(t)) * cosh(nu * t));
    }
    integral2 *= 0.5 * (t_max - 1.0);
    return integral + integral2;
}

double rc_gamma_fn(double x) {                           /* bessel.rs:100-123 */
    static const double COEF[9] = {
        0.9999999999998099, 676.5203681218851, -1259.1392167224028,
        771.3234287776531, -176.6150291621406, 12.507343278686905,
        -0.13857109526572012, 9.984369578019572e-6, 1.505632735149312e-7};
    if (x < 0.5) return M_PI / (sin(M_PI * x) * rc_gamma_fn(1.0 - x));
    double xm1 = x - 1.0;
    double a = COEF[0];
    double t = xm1 + 7.5;
    for (int i = 1; i < 9; i++) a += COEF[i] / (xm1 + (double)i);
    return sqrt(2.0 * M_PI) * pow(t, xm1 + 0.5) * exp(-t) * a;
}

/* ------------------------------------------------------------------ */
/* variogram/model.rs                                                  */
/* ------------------------------------------------------------------ */
double rc_shape_g(int32_t kind, double param, double h, double a) { /* model.rs:122-181 */
    double d = h / a;
    switch (kind) {
    case RC_SPHERICAL:
        return d >= 1.0 ? 1.0 : 1.5 * d - 0.5 * d * d * d;
    case RC_EXPONENTIAL:
        return 1.0 - exp(-d);
    case RC_GAUSSIAN:
        return 1.0 - exp(-(d * d));
    case RC_MATERN15: {
        double s = sqrt(3.0) * d;
        return 1.0 - (1.0 + s) * exp(-s);
    }
    case RC_MATERN25: {
        double s = sqrt(5.0) * d;
        return 1.0 - (1.0 + s + s * s / 3.0) * exp(-s);
    }
    case RC_MATERN: {
        double nu = param;
        if (d <= 0.0) return 0.0;
        double s = sqrt(2.0 * nu) * d;
        double corr = pow(2.0, 1.0 - nu) / rc_gamma_fn(nu) * pow(s, nu) * rc_bessel_k(nu, s);
        double cl = corr < 0.0 ? 0.0 : (corr > 1.0 ? 1.0 : corr);
        return 1.0 - cl;
    }
    case RC_CIRCULAR: {
        double dd = fmin(d, 1.0);
        return 1.0 - (2.0 / M_PI) * (acos(dd) - dd * sqrt(1.0 - dd * dd));
    }
    case RC_STABLE:
        return 1.0 - exp(-pow(d, param));
    case RC_HOLE:
        if (d <= 0.0) return 0.0;
        return 1.0 - sin(d) / d;
    case RC_POWER:
        return pow(fabs(h), param);
    case RC_WAVE: {
        if (d <= 0.0) return 0.0;
        double pd = M_PI * d;
        return 1.0 - sin(pd) / pd;
    }
    }
    return NAN;
}

static inline double to_radians(double deg) { return deg * (M_PI / 180.0); }

void rc_rotation_matrix_3d(const rc_aniso *a, double rot[9]) { /* model.rs:407-433 */
    double alpha = (a->azimuth_deg >= 0.0 && a->azimuth_deg < 270.0)
                       ? to_radians(90.0 - a->azimuth_deg)
                       : to_radians(450.0 - a->azimuth_deg);
    double beta = to_radians(-a->dip_deg);
    double theta = to_radians(a->rake_deg);
    double sina = sin(alpha), cosa = cos(alpha);
    double sinb = sin(beta), cosb = cos(beta);
    double sint = sin(theta), cost = cos(theta);
    double afac1 = 1.0 / a->ratio;
    double afac2 = 1.0 / a->ratio_z;
    rot[0] = cosb * cosa;
    rot[1] = cosb * sina;
    rot[2] = -sinb;
    rot[3] = afac1 * (-cost * sina + sint * sinb * cosa);
    rot[4] = afac1 * (cost * cosa + sint * sinb * sina);
    rot[5] = afac1 * (sint * cosb);
    rot[6] = afac2 * (sint * sina + cost * sinb * cosa);
    rot[7] = afac2 * (-sint * cosa + cost * sinb * sina);
    rot[8] = afac2 * (cost * cosb);
}

void rc_transform_vector(const rc_aniso *a, int32_t D, const double *dh, double *out) { /* model.rs:353-367 */
    if (D == 3) {
        double rot[9];
        rc_rotation_matrix_3d(a, rot);
        double h0 = dh[0], h1 = dh[1], h2 = dh[2];
        out[0] = rot[0] * h0 + rot[1] * h1 + rot[2] * h2;
        out[1] = rot[3] * h0 + rot[4] * h1 + rot[5] * h2;
        out[2] = rot[6] * h0 + rot[7] * h1 + rot[8] * h2;
    } else {
        double ang = to_radians(a->azimuth_deg);
        double s = sin(ang), c = cos(ang);
        out[0] = dh[0] * s + dh[1] * c;
        out[1] = (dh[0] * c - dh[1] * s) / a->ratio;
    }
}

static double effective_h(const rc_structure *s, int32_t D, const double *dh) { /* model.rs:521-539 */
    double acc = 0.0;
    if (!s->has_anis) {
        for (int d = 0; d < D; d++) acc += dh[d] * dh[d];
        return sqrt(acc);
    }
    double e[3];
    rc_transform_vector(&s->anis, D, dh, e);
    for (int d = 0; d < D; d++) acc += e[d] * e[d];
    return sqrt(acc);
}

double rc_gamma_dh(const rc_model *m, int32_t D, const double *dh) { /* model.rs:636-646 */
    int allzero = 1;
    for (int d = 0; d < D; d++)
        if (dh[d] != 0.0) allzero = 0;
    if (allzero) return 0.0;
    double sum = -0.0; /* Iterator::sum::<f64> left fold */
    for (int i = 0; i < m->n_structures; i++) {
        const rc_structure *s = &m->s[i];
        sum += s->sill * rc_shape_g(s->kind, s->param, effective_h(s, D, dh), s->range);
    }
    return m->nugget + sum;
}

double rc_total_sill(const rc_model *m) {                /* model.rs:658-660 */
    double sum = -0.0;
    for (int i = 0; i < m->n_structures; i++) sum += m->s[i].sill;
    return m->nugget + sum;
}

static int32_t aniso_validate(const rc_aniso *a) {       /* model.rs:376-400 */
    if (!(a->ratio_z > 0.0) || !isfinite(a->ratio_z)) return RC_INVALID_PARAMETER;
    if (!(a->ratio > 0.0) || !isfinite(a->ratio)) return RC_INVALID_PARAMETER;
    if (!isfinite(a->azimuth_deg)) return RC_INVALID_PARAMETER;
    if (!isfinite(a->dip_deg) || !isfinite(a->rake_deg)) return RC_INVALID_PARAMETER;
    return RC_OK;
}

int32_t rc_model_validate(const rc_model *m, int32_t D) { /* model.rs:554-618, 262-290 */
    if (!(m->nugget >= 0.0) || !isfinite(m->nugget)) return RC_INVALID_PARAMETER;
    if (m->n_structures < 0 || m->n_structures > RC_MAX_STRUCT) return RC_INVALID_PARAMETER;
    for (int i = 0; i < m->n_structures; i++) {
        const rc_structure *s = &m->s[i];
        if (!(s->sill >= 0.0) || !isfinite(s->sill)) return RC_INVALID_PARAMETER;
        if (!(s->range > 0.0) || !isfinite(s->range)) return RC_INVALID_PARAMETER;
        if (s->kind == RC_MATERN && (!(s->param > 0.0) || s->param > 15.0)) return RC_INVALID_PARAMETER;
        if (s->kind == RC_STABLE && (!(s->param > 0.0) || s->param > 2.0)) return RC_INVALID_PARAMETER;
        if (s->kind == RC_POWER && (!(s->param > 0.0) || s->param >= 2.0)) return RC_INVALID_PARAMETER;
        if (s->kind < 0 || s->kind > RC_POWER) return RC_INVALID_PARAMETER;
        if (s->kind == RC_CIRCULAR && D > 2) return RC_INVALID_PARAMETER;
        if (s->has_anis && aniso_validate(&s->anis) != RC_OK) return RC_INVALID_PARAMETER;
    }
    if (!(rc_total_sill(m) > 0.0)) return RC_INVALID_PARAMETER;
    return RC_OK;
}

static int has_power(const rc_model *m) {
    for (int i = 0; i < m->n_structures; i++)
        if (m->s[i].kind == RC_POWER) return 1;
    return 0;
}

/* ------------------------------------------------------------------ */
/* linalg.rs                                                           */
/* ------------------------------------------------------------------ */
/* lu_factor + Lu::solve (linalg.rs:27-111): row-major a[n*n], b[n]. */
int32_t rc_lu_solve(double *a, int32_t n, double *b) {
    double scale = 0.0;
    for (int64_t i = 0; i < (int64_t)n * n; i++) scale = fmax(scale, fabs(a[i]));
    scale = fmax(scale, 2.2250738585072014e-308);
    int32_t *piv = (int32_t *)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
    for (int k = 0; k < n; k++) {
        int p = k;
        double mx = fabs(a[k * n + k]);
        for (int i = k + 1; i < n; i++) {
            double v = fabs(a[i * n + k]);
            if (v > mx) { mx = v; p = i; }
        }
        if (mx < 1e-12 * scale) { free(piv); return RC_SINGULAR_SYSTEM; }
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < n; j++) {
                double t = a[k * n + j]; a[k * n + j] = a[p * n + j]; a[p * n + j] = t;
            }
        double pivot = a[k * n + k];
        for (int i = k + 1; i < n; i++) {
            double f = a[i * n + k] / pivot;
            a[i * n + k] = f;
            if (f != 0.0)
                for (int j = k + 1; j < n; j++) a[i * n + j] -= f * a[k * n + j];
        }
    }
    for (int k = 0; k < n; k++) {
        double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t;
    }
    for (int i = 0; i < n; i++) {
        double s = b[i];
        for (int j = 0; j < i; j++) s -= a[i * n + j] * b[j];
        b[i] = s;
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int j = i + 1; j < n; j++) s -= a[i * n + j] * b[j];
        b[i] = s / a[i * n + i];
    }
    free(piv);
    return RC_OK;
}

/* ln|det| from an LU factor is needed by the dense-MVN test oracle
 * (vecchia.rs:1958-1974 uses lu_factor + ln_det_abs, linalg.rs:80-84). */
static int32_t lu_solve_logdet(double *a, int32_t n, double *b, double *logdet) {
    int32_t st = rc_lu_solve(a, n, b);
    if (st != RC_OK) return st;
    double s = -0.0;
    for (int i = 0; i < n; i++) s += log(fabs(a[i * n + i]));
    *logdet = s;
    return RC_OK;
}

/* cholesky_factor_in_place / forward / back (linalg.rs:123-203). */
int32_t rc_cholesky_solve(double *a, int32_t n, double *b) {
    double scale = 0.0;
    for (int i = 0; i < n; i++) scale = fmax(scale, a[i * n + i]);
    scale = fmax(scale, 2.2250738585072014e-308);
    for (int k = 0; k < n; k++) {
        double d = a[k * n + k];
        for (int j = 0; j < k; j++) d -= a[k * n + j] * a[k * n + j];
        if (!(d > 1e-12 * scale)) return RC_SINGULAR_SYSTEM;
        d = sqrt(d);
        a[k * n + k] = d;
        for (int i = k + 1; i < n; i++) {
            double s = a[i * n + k];
            for (int j = 0; j < k; j++) s -= a[i * n + j] * a[k * n + j];
            a[i * n + k] = s / d;
        }
    }
    for (int i = 0; i < n; i++) {
        double s = b[i];
        for (int j = 0; j < i; j++) s -= a[i * n + j] * b[j];
        b[i] = s / a[i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int j = i + 1; j < n; j++) s -= a[j * n + i] * b[j];
        b[i] = s / a[i * n + i];
    }
    return RC_OK;
}

/* ------------------------------------------------------------------ */
/* data.rs:150-163 — first duplicate pair on (c + 0.0).to_bits().       */
/* ------------------------------------------------------------------ */
typedef struct { uint64_t k[3]; int64_t i; } dupkey;
static int g_dup_D;
static int dupcmp(const void *pa, const void *pb) {
    const dupkey *a = (const dupkey *)pa, *b = (const dupkey *)pb;
    for (int d = 0; d < g_dup_D; d++) {
        if (a->k[d] < b->k[d]) return -1;
        if (a->k[d] > b->k[d]) return 1;
    }
    return (a->i < b->i) ? -1 : (a->i > b->i);
}
static pthread_mutex_t dup_mu = PTHREAD_MUTEX_INITIALIZER;

int32_t rc_duplicate_pair(int32_t D, const double *coords, int64_t n, int64_t *pair) {
    dupkey *v = (dupkey *)malloc(sizeof(dupkey) * (n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) {
        for (int d = 0; d < 3; d++) v[i].k[d] = 0;
        for (int d = 0; d < D; d++) {
            double c = coords[i * D + d] + 0.0;
            memcpy(&v[i].k[d], &c, 8);
        }
        v[i].i = i;
    }
    pthread_mutex_lock(&dup_mu);
    g_dup_D = D;
    qsort(v, (size_t)n, sizeof(dupkey), dupcmp);
    pthread_mutex_unlock(&dup_mu);
    /* HashMap scan in index order returns (first occurrence, i) for the
     * smallest i whose key was already seen. */
    int64_t best_i = -1, best_j = -1;
    for (int64_t s = 0; s + 1 < n; s++) {
        int same = 1;
        for (int d = 0; d < D; d++)
            if (v[s].k[d] != v[s + 1].k[d]) same = 0;
        if (same && (s == 0 || memcmp(v[s - 1].k, v[s].k, sizeof(uint64_t) * 3) != 0)) {
            if (best_j < 0 || v[s + 1].i < best_j) { best_i = v[s].i; best_j = v[s + 1].i; }
        }
    }
    free(v);
    if (best_j >= 0) { pair[0] = best_i; pair[1] = best_j; return 1; }
    return 0;
}

/* ------------------------------------------------------------------ */
/* search.rs — KdTree (static, median split) + BucketGrid + heap rules */
/* ------------------------------------------------------------------ */
typedef struct { uint64_t d2; uint32_t idx; } hent;
typedef struct { hent *v; int64_t len; int64_t cap; } mheap;

static inline int hgt(hent a, hent b) { return a.d2 > b.d2 || (a.d2 == b.d2 && a.idx > b.idx); }
static void heap_push(mheap *h, hent e) {
    int64_t i = h->len++;
    h->v[i] = e;
    while (i > 0) {
        int64_t p = (i - 1) / 2;
        if (hgt(h->v[i], h->v[p])) { hent t = h->v[i]; h->v[i] = h->v[p]; h->v[p] = t; i = p; }
        else break;
    }
}
static void heap_replace_top(mheap *h, hent e) {
    int64_t i = 0;
    h->v[0] = e;
    for (;;) {
        int64_t l = 2 * i + 1, r = l + 1, g = i;
        if (l < h->len && hgt(h->v[l], h->v[g])) g = l;
        if (r < h->len && hgt(h->v[r], h->v[g])) g = r;
        if (g == i) break;
        hent t = h->v[i]; h->v[i] = h->v[g]; h->v[g] = t; i = g;
    }
}
static inline uint64_t dbits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double bitsd(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

/* push_candidate (search.rs:15-22): full heap replaces only if d2 < worst
 * (strict, on d2 alone); the removed entry is the max (d2, idx). */
static void push_candidate(mheap *h, int64_t k, double d2, uint32_t idx) {
    hent e = {dbits(d2), idx};
    if (h->len < k) heap_push(h, e);
    else if (e.d2 < h->v[0].d2) heap_replace_top(h, e);
}
static double heap_worst(const mheap *h, int64_t k) {   /* search.rs:24-30 */
    return h->len == k ? bitsd(h->v[0].d2) : INFINITY;
}
static int hent_cmp(const void *pa, const void *pb) {   /* into_sorted: search.rs:32-36 */
    const hent *a = (const hent *)pa, *b = (const hent *)pb;
    if (a->d2 != b->d2) return a->d2 < b->d2 ? -1 : 1;
    return (a->idx < b->idx) ? -1 : (a->idx > b->idx);
}

/* total_cmp key for f64 (IEEE totalOrder as signed integer). */
static inline int64_t total_key(double x) {
    int64_t b;
    memcpy(&b, &x, 8);
    return b ^ (int64_t)(((uint64_t)(b >> 63)) >> 1);
}

typedef struct {
    int32_t D;
    int64_t n;
    uint32_t *idx;       /* tree layout (permutation) */
    const double *coords; /* [n][D] in original index order */
} kdtree;

typedef struct { int64_t key; uint32_t i; } selkey;
static int selcmp(const void *pa, const void *pb) {
    const selkey *a = (const selkey *)pa, *b = (const selkey *)pb;
    if (a->key != b->key) return a->key < b->key ? -1 : 1;
    return (a->i < b->i) ? -1 : (a->i > b->i);
}

/* build_rec (search.rs:102-113).  Rust's select_nth_unstable_by(total_cmp)
 * leaves the slice partitioned around the median; with distinct axis
 * values the resulting tree is unique.  Ties on the axis are broken by
 * index here (parity unpinned for duplicate axis coordinates: the
 * reference's placement then depends on its selection algorithm). */
static void build_rec(uint32_t *idx, int64_t len, const double *coords, int32_t D, int32_t axis,
                      selkey *tmp) {
    if (len <= 1) return;
    int64_t mid = len / 2;
    for (int64_t i = 0; i < len; i++) {
        tmp[i].key = total_key(coords[(int64_t)idx[i] * D + axis]);
        tmp[i].i = idx[i];
    }
    qsort(tmp, (size_t)len, sizeof(selkey), selcmp);
    for (int64_t i = 0; i < len; i++) idx[i] = tmp[i].i;
    build_rec(idx, mid, coords, D, (axis + 1) % D, tmp);
    build_rec(idx + mid + 1, len - mid - 1, coords, D, (axis + 1) % D, tmp);
}

static void kd_build(kdtree *t, const double *coords, int64_t n, int32_t D) { /* search.rs:47-54 */
    t->D = D;
    t->n = n;
    t->coords = coords;
    t->idx = (uint32_t *)malloc(sizeof(uint32_t) * (n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) t->idx[i] = (uint32_t)i;
    selkey *tmp = (selkey *)malloc(sizeof(selkey) * (n > 0 ? n : 1));
    build_rec(t->idx, n, coords, D, 0, tmp);
    free(tmp);
}

int32_t rc_kdtree_layout(int32_t D, const double *coords, int64_t n, uint32_t *idx_out) {
    kdtree t;
    kd_build(&t, coords, n, D);
    memcpy(idx_out, t.idx, sizeof(uint32_t) * n);
    free(t.idx);
    return RC_OK;
}

static void kd_search(const kdtree *t, const uint32_t *node, int64_t len, int32_t axis,
                      const double *target, int64_t k, double r2, mheap *h) { /* search.rs:67-99 */
    if (len == 0) return;
    int32_t D = t->D;
    int64_t mid = len / 2;
    const double *p = t->coords + (int64_t)node[mid] * D;
    double d2 = 0.0;
    for (int d = 0; d < D; d++) {
        double dd = p[d] - target[d];
        d2 += dd * dd;
    }
    if (d2 <= r2) push_candidate(h, k, d2, node[mid]);
    double diff = target[axis] - p[axis];
    const uint32_t *nearp, *farp;
    int64_t nearl, farl;
    if (diff <= 0.0) { nearp = node; nearl = mid; farp = node + mid + 1; farl = len - mid - 1; }
    else { nearp = node + mid + 1; nearl = len - mid - 1; farp = node; farl = mid; }
    kd_search(t, nearp, nearl, (axis + 1) % D, target, k, r2, h);
    if (diff * diff <= fmin(heap_worst(h, k), r2))
        kd_search(t, farp, farl, (axis + 1) % D, target, k, r2, h);
}

/* KdTree::k_nearest (search.rs:57-65): returns count, writes idx sorted. */
static int64_t kd_knn(const kdtree *t, const double *target, int64_t k, double r2, mheap *h,
                      uint32_t *out) {
    if (k == 0 || t->n == 0) return 0;
    h->len = 0;
    kd_search(t, t->idx, t->n, 0, target, k, r2, h);
    qsort(h->v, (size_t)h->len, sizeof(hent), hent_cmp);
    for (int64_t i = 0; i < h->len; i++) out[i] = h->v[i].idx;
    return h->len;
}

int64_t rc_kdtree_knn(int32_t D, const double *coords, int64_t n, const double *targets, int64_t T,
                      int64_t k, int32_t has_radius, double radius, int32_t *out_idx,
                      int32_t *out_count) {
    kdtree t;
    kd_build(&t, coords, n, D);
    mheap h;
    h.cap = k + 1;
    h.v = (hent *)malloc(sizeof(hent) * (size_t)(k + 1));
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(k + 1));
    double r2 = has_radius ? radius * radius : INFINITY;
    for (int64_t ti = 0; ti < T; ti++) {
        int64_t c = kd_knn(&t, targets + ti * D, k, r2, &h, tmp);
        out_count[ti] = (int32_t)c;
        for (int64_t j = 0; j < k; j++) out_idx[ti * k + j] = j < c ? (int32_t)tmp[j] : -1;
    }
    free(tmp);
    free(h.v);
    free(t.idx);
    return T;
}

/* BucketGrid (search.rs:119-313) */
typedef struct {
    int32_t D;
    double origin[3];
    double cell;
    int64_t nn[3];
    int64_t strides[3];
    int64_t nbuckets;
    int64_t *bstart;   /* bucket contents as linked chunks: simple vectors */
    uint32_t **bv;
    int64_t *blen, *bcap;
    double *pts;       /* [cap][D] insertion order */
    int64_t npts, ptscap;
} bgrid;

static void bgrid_new(bgrid *g, int32_t D, const double *min, const double *max, int64_t expected) { /* search.rs:138-197 */
    g->D = D;
    double ext[3];
    double max_extent = 0.0;
    for (int d = 0; d < D; d++) { ext[d] = max[d] - min[d]; max_extent = fmax(max_extent, ext[d]); }
    double tol = max_extent * 1e-9;
    int active[3];
    int n_active = 0;
    for (int d = 0; d < D; d++) { active[d] = ext[d] > tol; n_active += active[d]; }
    if (n_active < 1) n_active = 1;
    double volume = 1.0;
    for (int d = 0; d < D; d++) if (active[d]) volume *= ext[d];
    int64_t em = expected > 1 ? expected : 1;
    double cell = fmax(pow(volume, 1.0 / (double)n_active), 1e-12) *
                  (1.0 / pow((double)em, 1.0 / (double)n_active));
    int64_t cap = em * 4; /* saturating_mul is irrelevant at these sizes */
    if (cap < 64) cap = 64;
    for (;;) {
        int64_t total = 1;
        int overflowed = 0;
        for (int d = 0; d < D; d++) {
            if (active[d]) {
                double q = ceil(ext[d] / cell);
                int64_t v = (q >= 9.2e18) ? INT64_MAX : (q <= 0.0 || q != q ? 0 : (int64_t)q);
                g->nn[d] = v > 1 ? v : 1;
            } else g->nn[d] = 1;
            if (total > INT64_MAX / g->nn[d]) { overflowed = 1; break; }
            total *= g->nn[d];
        }
        if (!overflowed && total <= cap) break;
        cell *= 2.0;
    }
    int64_t total = 1;
    for (int d = 0; d < D; d++) { g->strides[d] = total; total *= g->nn[d]; g->origin[d] = min[d]; }
    g->cell = cell;
    g->nbuckets = total;
    g->bv = (uint32_t **)calloc((size_t)total, sizeof(uint32_t *));
    g->blen = (int64_t *)calloc((size_t)total, sizeof(int64_t));
    g->bcap = (int64_t *)calloc((size_t)total, sizeof(int64_t));
    g->ptscap = em;
    g->pts = (double *)malloc(sizeof(double) * (size_t)(em * D));
    g->npts = 0;
}
static void bgrid_free(bgrid *g) {
    for (int64_t b = 0; b < g->nbuckets; b++) free(g->bv[b]);
    free(g->bv); free(g->blen); free(g->bcap); free(g->pts);
}
/* `(x / cell) as isize` saturates and truncates; NaN -> 0. */
static inline int64_t sat_isize(double v) {
    if (v != v) return 0;
    if (v >= 9.2233720368547758e18) return INT64_MAX;
    if (v <= -9.2233720368547758e18) return INT64_MIN;
    return (int64_t)v;
}
static void bgrid_cell_of(const bgrid *g, const double *p, int64_t *c) { /* search.rs:199-206 */
    for (int d = 0; d < g->D; d++) {
        int64_t v = sat_isize((p[d] - g->origin[d]) / g->cell);
        int64_t hi = g->nn[d] - 1;
        c[d] = v < 0 ? 0 : (v > hi ? hi : v);
    }
}
static void bgrid_insert(bgrid *g, const double *p) {     /* search.rs:213-218 */
    int64_t c[3];
    bgrid_cell_of(g, p, c);
    int64_t b = 0;
    for (int d = 0; d < g->D; d++) b += c[d] * g->strides[d];
    if (g->blen[b] == g->bcap[b]) {
        g->bcap[b] = g->bcap[b] ? g->bcap[b] * 2 : 4;
        g->bv[b] = (uint32_t *)realloc(g->bv[b], sizeof(uint32_t) * (size_t)g->bcap[b]);
    }
    g->bv[b][g->blen[b]++] = (uint32_t)g->npts;
    if (g->npts == g->ptscap) {
        g->ptscap *= 2;
        g->pts = (double *)realloc(g->pts, sizeof(double) * (size_t)(g->ptscap * g->D));
    }
    for (int d = 0; d < g->D; d++) g->pts[g->npts * g->D + d] = p[d];
    g->npts++;
}
static void bgrid_visit_bucket(const bgrid *g, const int64_t *cur, const double *target, int64_t k,
                               double r2, mheap *h) {
    for (int d = 0; d < g->D; d++)
        if (cur[d] < 0 || cur[d] >= g->nn[d]) return;
    int64_t b = 0;
    for (int d = 0; d < g->D; d++) b += cur[d] * g->strides[d];
    for (int64_t q = 0; q < g->blen[b]; q++) {
        uint32_t i = g->bv[b][q];
        const double *p = g->pts + (int64_t)i * g->D;
        double d2 = 0.0;
        for (int d = 0; d < g->D; d++) {
            double dd = p[d] - target[d];
            d2 += dd * dd;
        }
        if (d2 <= r2) push_candidate(h, k, d2, i);
    }
}
/* for_shell_cells (search.rs:256-313): faces of each dimension, odometer
 * with dimension 0 fastest. */
static void bgrid_shell(const bgrid *g, const int64_t *tc, int64_t ring, const double *target,
                        int64_t k, double r2, mheap *h) {
    int D = g->D;
    if (ring == 0) { bgrid_visit_bucket(g, tc, target, k, r2, h); return; }
    int64_t r = ring;
    for (int fd = 0; fd < D; fd++) {
        for (int sd = 0; sd < 2; sd++) {
            int64_t side = sd == 0 ? -r : r;
            int64_t lo[3], hi[3];
            int empty = 0;
            for (int d = 0; d < D; d++) {
                if (d < fd) { lo[d] = tc[d] - r + 1; hi[d] = tc[d] + r - 1; }
                else if (d == fd) { lo[d] = tc[d] + side; hi[d] = tc[d] + side; }
                else { lo[d] = tc[d] - r; hi[d] = tc[d] + r; }
                if (lo[d] > hi[d]) empty = 1;
            }
            if (empty) continue;
            int64_t cur[3];
            for (int d = 0; d < D; d++) cur[d] = lo[d];
            for (;;) {
                bgrid_visit_bucket(g, cur, target, k, r2, h);
                int d = 0;
                for (; d < D; d++) {
                    cur[d] += 1;
                    if (cur[d] <= hi[d]) break;
                    cur[d] = lo[d];
                }
                if (d == D) break;
            }
        }
    }
}
/* BucketGrid::k_nearest (search.rs:221-251): count, sorted positions. */
static int64_t bgrid_knn(const bgrid *g, const double *target, int64_t k, double r2, mheap *h,
                         uint32_t *out) {
    if (k == 0 || g->npts == 0) return 0;
    h->len = 0;
    int64_t tc[3];
    bgrid_cell_of(g, target, tc);
    int64_t max_ring = g->nn[0];
    for (int d = 1; d < g->D; d++) if (g->nn[d] > max_ring) max_ring = g->nn[d];
    for (int64_t ring = 0; ring <= max_ring; ring++) {
        double rm = fmax((double)ring - 1.0, 0.0) * g->cell;
        if (rm * rm > fmin(heap_worst(h, k), r2)) break;
        bgrid_shell(g, tc, ring, target, k, r2, h);
    }
    qsort(h->v, (size_t)h->len, sizeof(hent), hent_cmp);
    for (int64_t i = 0; i < h->len; i++) out[i] = h->v[i].idx;
    return h->len;
}

/* ------------------------------------------------------------------ */
/* kriging.rs                                                          */
/* ------------------------------------------------------------------ */
typedef struct {
    int32_t D;
    int64_t n;
    const double *coords, *values;
    const rc_model *model;
    rc_kconfig cfg;
    int has_tree;
    kdtree tree;
    double *tcoords;   /* search-frame coords (kriging.rs:375-380) */
    double center[3], scale;
    int32_t m;
    int power;
    double c0;
    const double *err;          /* measurement-error variances (n) or NULL */
    int32_t nv;                 /* external covariates */
    const double *drift;        /* n x nv raw covariates or NULL */
    double dmean[8], dscale[8]; /* per-covariate normalization (kriging.rs:349-364) */
} krig;

static int32_t n_drift(const rc_kconfig *c, int32_t D) { /* kriging.rs:50-59 */
    switch (c->method) {
    case RC_SIMPLE: return 0;
    case RC_ORDINARY: return 1;
    case RC_UNIVERSAL:
        if (c->degree == 1) return 1 + D;
        if (c->degree == 2) return 1 + D + D * (D + 1) / 2;
        return 0;
    case RC_EXTERNAL_DRIFT: return 1 + c->degree;   /* degree holds n_vars */
    }
    return 0;
}

static void search_point(const krig *K, const double *p, double *out) { /* kriging.rs:16-21 */
    if (K->cfg.has_anis_search) rc_transform_vector(&K->cfg.anis_search, K->D, p, out);
    else for (int d = 0; d < K->D; d++) out[d] = p[d];
}

/* Kriging::build (kriging.rs:262-408) — validation, centre/scale, tree. */
static int32_t krig_build(krig *K, int32_t D, const double *coords, const double *values, int64_t n,
                          const rc_model *model, const rc_kconfig *cfg, int64_t *dup_pair) {
    memset(K, 0, sizeof(*K));
    K->D = D; K->n = n; K->coords = coords; K->values = values; K->model = model; K->cfg = *cfg;
    if (n <= 0) return RC_INSUFFICIENT_DATA;                 /* data.rs:58-60 */
    for (int64_t i = 0; i < n * D; i++) if (!isfinite(coords[i])) return RC_INVALID_PARAMETER;
    for (int64_t i = 0; i < n; i++) if (!isfinite(values[i])) return RC_INVALID_PARAMETER;
    if (rc_model_validate(model, D) != RC_OK) return RC_INVALID_PARAMETER;
    if (cfg->method == RC_UNIVERSAL && !(cfg->degree == 1 || cfg->degree == 2)) return RC_INVALID_PARAMETER;
    K->power = has_power(model);
    if (K->power && cfg->method == RC_SIMPLE) return RC_INVALID_PARAMETER;
    if (cfg->has_radius && !(cfg->radius > 0.0)) return RC_INVALID_PARAMETER;
    if (cfg->has_anis_search && aniso_validate(&cfg->anis_search) != RC_OK) return RC_INVALID_PARAMETER;
    if (cfg->max_neighbors == 0) return RC_INVALID_PARAMETER;
    if (rc_duplicate_pair(D, coords, n, dup_pair)) return RC_DUPLICATE_POINTS;
    if (cfg->max_per_octant == 0) return RC_INVALID_PARAMETER;
    if (cfg->min_neighbors >= 0) {
        if (cfg->min_neighbors == 0) return RC_INVALID_PARAMETER;
        if (cfg->min_neighbors > n) return RC_INSUFFICIENT_DATA;
        if (cfg->max_neighbors >= 0 && cfg->min_neighbors > cfg->max_neighbors) return RC_INVALID_PARAMETER;
    }
    double mn[3], mx[3];
    for (int d = 0; d < D; d++) { mn[d] = INFINITY; mx[d] = -INFINITY; }
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < D; d++) {
            mn[d] = fmin(mn[d], coords[i * D + d]);
            mx[d] = fmax(mx[d], coords[i * D + d]);
        }
    double spread = 0.0;
    for (int d = 0; d < D; d++) {
        K->center[d] = 0.5 * (mn[d] + mx[d]);
        spread = fmax(spread, mx[d] - mn[d]);
    }
    K->scale = fmax(0.5 * spread, 2.2250738585072014e-308);
    K->m = n_drift(cfg, D);
    K->c0 = rc_total_sill(model);
    K->nv = 0;
    K->has_tree = cfg->max_neighbors >= 0 || cfg->has_radius || cfg->max_per_octant >= 0;
    if (K->has_tree) {
        K->tcoords = (double *)malloc(sizeof(double) * (size_t)(n * D));
        for (int64_t i = 0; i < n; i++) search_point(K, coords + i * D, K->tcoords + i * D);
        kd_build(&K->tree, K->tcoords, n, D);
    }
    return RC_OK;
}
static void krig_free(krig *K) {
    if (K->has_tree) { free(K->tcoords); free(K->tree.idx); }
}

static void fill_basis(const krig *K, const double *p, const double *drift, double *out) { /* kriging.rs:410-439 */
    int D = K->D;
    switch (K->cfg.method) {
    case RC_EXTERNAL_DRIFT:
        out[0] = 1.0;
        for (int k = 0; k < K->nv; k++) out[1 + k] = (drift[k] - K->dmean[k]) / K->dscale[k];
        break;
    case RC_SIMPLE: break;
    case RC_ORDINARY: out[0] = 1.0; break;
    case RC_UNIVERSAL: {
        double xn[3];
        for (int d = 0; d < D; d++) xn[d] = (p[d] - K->center[d]) / K->scale;
        out[0] = 1.0;
        for (int d = 0; d < D; d++) out[1 + d] = xn[d];
        if (K->cfg.degree == 2) {
            int idx = 1 + D;
            for (int a = 0; a < D; a++)
                for (int b = a; b < D; b++) out[idx++] = xn[a] * xn[b];
        }
    } break;
    }
}

typedef struct {
    mheap h;
    uint32_t *cand;
    int64_t *nb;
    double *A, *b, *b0, *f;
} kws;

static void kws_init(kws *w, const krig *K) {
    int64_t n = K->n;
    w->h.cap = n + 1;
    w->h.v = (hent *)malloc(sizeof(hent) * (size_t)(n + 1));
    w->cand = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n + 1));
    w->nb = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    w->A = NULL; w->b = NULL; w->b0 = NULL; w->f = NULL;
}
static void kws_free(kws *w) {
    free(w->h.v); free(w->cand); free(w->nb); free(w->A); free(w->b); free(w->b0); free(w->f);
}

/* Kriging::neighbors (kriging.rs:441-490). */
static int64_t krig_neighbors(const krig *K, const double *target, kws *w) {
    int D = K->D;
    int64_t n = K->n;
    if (!K->has_tree) {
        for (int64_t i = 0; i < n; i++) w->nb[i] = i;
        return n;
    }
    double st[3];
    search_point(K, target, st);
    double r2 = K->cfg.has_radius ? K->cfg.radius * K->cfg.radius : INFINITY;
    if (K->cfg.max_per_octant < 0) {
        int64_t k = K->cfg.max_neighbors >= 0 ? K->cfg.max_neighbors : n;
        int64_t c = kd_knn(&K->tree, st, k, r2, &w->h, w->cand);
        for (int64_t i = 0; i < c; i++) w->nb[i] = w->cand[i];
        return c;
    }
    int64_t c = kd_knn(&K->tree, st, n, r2, &w->h, w->cand);
    /* The stable re-sort by recomputed d2 (kriging.rs:466-470) recomputes
     * exactly the d2 the heap ordered by (same transformed coordinates,
     * same left-to-right sum), so the (d2, idx)-sorted list is unchanged. */
    int64_t counts[8] = {0};
    int64_t sel = 0;
    for (int64_t q = 0; q < c; q++) {
        uint32_t i = w->cand[q];
        const double *cc = K->tcoords + (int64_t)i * D;
        int oct = 0;
        for (int d = 0; d < D; d++) if (cc[d] >= st[d]) oct |= 1 << d;
        if (counts[oct] < K->cfg.max_per_octant) { counts[oct]++; w->nb[sel++] = i; }
    }
    if (K->cfg.max_neighbors >= 0 && sel > K->cfg.max_neighbors) sel = K->cfg.max_neighbors;
    return sel;
}

static inline double cov_entry(const krig *K, const double *dh) {
    double g = rc_gamma_dh(K->model, K->D, dh);
    return K->power ? g : K->c0 - g;
}

/* build_lhs (kriging.rs:527-567) + predict_inner (569-644); tdrift: the
 * target's covariates (external drift). */
static int32_t krig_predict(const krig *K, const double *target, const double *tdrift, kws *w,
                            double *val, double *var, double *lag, int *has_lag) {
    int D = K->D;
    int64_t nn = krig_neighbors(K, target, w);
    int64_t min_nb = K->cfg.min_neighbors >= 0 ? K->cfg.min_neighbors : 1;
    if (nn < min_nb) return RC_NO_NEIGHBORS;
    int64_t m = K->m;
    if (nn < m) return RC_INSUFFICIENT_DATA;
    int64_t dim = nn + m;
    w->A = (double *)realloc(w->A, sizeof(double) * (size_t)(dim * dim));
    w->b = (double *)realloc(w->b, sizeof(double) * (size_t)dim);
    w->b0 = (double *)realloc(w->b0, sizeof(double) * (size_t)dim);
    w->f = (double *)realloc(w->f, sizeof(double) * (size_t)(m > 0 ? m : 1));
    double *A = w->A, *b = w->b, *f = w->f;
    double zero[3] = {0.0, 0.0, 0.0};
    for (int64_t ii = 0; ii < nn; ii++) {
        const double *pi = K->coords + w->nb[ii] * D;
        double dh[3];
        for (int d = 0; d < D; d++) dh[d] = pi[d] - target[d];
        b[ii] = cov_entry(K, dh);
    }
    fill_basis(K, target, tdrift, f);
    for (int64_t q = 0; q < m; q++) b[nn + q] = f[q];
    memcpy(w->b0, b, sizeof(double) * (size_t)dim);
    memset(A, 0, sizeof(double) * (size_t)(dim * dim));
    for (int64_t ii = 0; ii < nn; ii++) {
        const double *pi = K->coords + w->nb[ii] * D;
        A[ii * dim + ii] = cov_entry(K, zero) + (K->err ? K->err[w->nb[ii]] : 0.0);
        for (int64_t jj = ii + 1; jj < nn; jj++) {
            const double *pj = K->coords + w->nb[jj] * D;
            double dh[3];
            for (int d = 0; d < D; d++) dh[d] = pi[d] - pj[d];
            double c = cov_entry(K, dh);
            A[ii * dim + jj] = c;
            A[jj * dim + ii] = c;
        }
        fill_basis(K, pi, K->drift ? K->drift + w->nb[ii] * K->nv : NULL, f);
        for (int64_t q = 0; q < m; q++) {
            A[ii * dim + nn + q] = f[q];
            A[(nn + q) * dim + ii] = f[q];
        }
    }
    int32_t st = rc_lu_solve(A, (int32_t)dim, b);
    if (st != RC_OK) return st;
    const double *x = b, *b0 = w->b0;
    if (K->cfg.method == RC_SIMPLE) {
        double mean = K->cfg.simple_mean;
        double v = mean, red = 0.0;
        for (int64_t ii = 0; ii < nn; ii++) {
            v += x[ii] * (K->values[w->nb[ii]] - mean);
            red += x[ii] * b0[ii];
        }
        *val = v;
        *var = fmax(K->c0 - red, 0.0);
        *has_lag = 0;
        *lag = NAN;
    } else {
        double v = -0.0, red = -0.0;
        for (int64_t ii = 0; ii < nn; ii++) v += x[ii] * K->values[w->nb[ii]];
        for (int64_t i = 0; i < dim; i++) red += x[i] * b0[i];
        double variance = K->power ? red : K->c0 - red;
        *val = v;
        *var = fmax(variance, 0.0);
        *has_lag = 1;
        *lag = x[nn];
    }
    return RC_OK;
}

/* predict_block (kriging.rs:678-777): block average of centre + offsets[u]. */
static int32_t krig_predict_block(const krig *K, const double *center, const double *offs,
                                  int64_t n_off, kws *w, double *val, double *var, double *lag,
                                  int *has_lag) {
    int D = K->D;
    if (!(K->cfg.method == RC_SIMPLE || K->cfg.method == RC_ORDINARY)) return RC_INVALID_PARAMETER;
    if (K->power) return RC_INVALID_PARAMETER;
    if (n_off == 0) return RC_INVALID_PARAMETER;
    int64_t nn = krig_neighbors(K, center, w);
    int64_t min_nb = K->cfg.min_neighbors >= 0 ? K->cfg.min_neighbors : 1;
    if (nn < min_nb) return RC_NO_NEIGHBORS;
    int64_t m = K->m;
    int64_t dim = nn + m;
    double c0 = K->c0;
    double nu = (double)n_off;
    w->A = (double *)realloc(w->A, sizeof(double) * (size_t)(dim * dim));
    w->b = (double *)realloc(w->b, sizeof(double) * (size_t)dim);
    w->b0 = (double *)realloc(w->b0, sizeof(double) * (size_t)dim);
    double *A = w->A, *b = w->b;
    memset(A, 0, sizeof(double) * (size_t)(dim * dim));
    for (int64_t ii = 0; ii < nn; ii++) {
        const double *pi = K->coords + w->nb[ii] * D;
        A[ii * dim + ii] = c0 + (K->err ? K->err[w->nb[ii]] : 0.0);
        for (int64_t jj = ii + 1; jj < nn; jj++) {
            const double *pj = K->coords + w->nb[jj] * D;
            double dh[3];
            for (int d = 0; d < D; d++) dh[d] = pi[d] - pj[d];
            double c = c0 - rc_gamma_dh(K->model, D, dh);
            A[ii * dim + jj] = c;
            A[jj * dim + ii] = c;
        }
        if (m == 1) {
            A[ii * dim + nn] = 1.0;
            A[nn * dim + ii] = 1.0;
        }
        double acc = -0.0;
        for (int64_t u = 0; u < n_off; u++) {
            double dh[3];
            for (int d = 0; d < D; d++) dh[d] = pi[d] - (center[d] + offs[u * D + d]);
            acc += c0 - rc_gamma_dh(K->model, D, dh);
        }
        b[ii] = acc / nu;
    }
    if (m == 1) b[nn] = 1.0;
    double c0c = c0 - K->model->nugget;
    double cbb = 0.0;
    for (int64_t u = 0; u < n_off; u++)
        for (int64_t v = 0; v < n_off; v++) {
            if (u == v) cbb += c0c;
            else {
                double dh[3];
                for (int d = 0; d < D; d++) dh[d] = offs[u * D + d] - offs[v * D + d];
                cbb += c0 - rc_gamma_dh(K->model, D, dh);
            }
        }
    cbb /= nu * nu;
    memcpy(w->b0, b, sizeof(double) * (size_t)dim);
    int32_t st = rc_lu_solve(A, (int32_t)dim, b);
    if (st != RC_OK) return st;
    const double *x = b, *b0 = w->b0;
    if (K->cfg.method == RC_SIMPLE) {
        double mean = K->cfg.simple_mean, v = mean, red = 0.0;
        for (int64_t ii = 0; ii < nn; ii++) {
            v += x[ii] * (K->values[w->nb[ii]] - mean);
            red += x[ii] * b0[ii];
        }
        *val = v;
        *var = fmax(cbb - red, 0.0);
        *has_lag = 0;
        *lag = NAN;
    } else {
        double v = -0.0, red = -0.0;
        for (int64_t ii = 0; ii < nn; ii++) v += x[ii] * K->values[w->nb[ii]];
        for (int64_t i = 0; i < dim; i++) red += x[i] * b0[i];
        *val = v;
        *var = fmax(cbb - red, 0.0);
        *has_lag = 1;
        *lag = x[nn];
    }
    return RC_OK;
}

/* parallel.rs:9-17 analogue: dynamic chunks over targets, output in order. */
typedef struct {
    const krig *K;
    const double *targets;
    const double *tdrift;     /* T x nv target covariates (predict_with_drift) */
    const double *offsets;    /* block discretization (predict_block) */
    int64_t n_off;
    int block;
    int64_t T;
    atomic_llong next;
    double *value, *variance, *lagrange;
    int32_t *status;
    /* neighbour mode */
    int64_t max_out;
    int32_t *idx_out, *count_out;
} kjob;

static void *kworker(void *arg) {
    kjob *J = (kjob *)arg;
    kws w;
    kws_init(&w, J->K);
    const int64_t CH = 64;
    for (;;) {
        int64_t s = atomic_fetch_add(&J->next, CH);
        if (s >= J->T) break;
        int64_t e = s + CH < J->T ? s + CH : J->T;
        for (int64_t t = s; t < e; t++) {
            const double *tg = J->targets + t * J->K->D;
            if (J->idx_out) {
                int64_t c = krig_neighbors(J->K, tg, &w);
                J->count_out[t] = (int32_t)c;
                for (int64_t q = 0; q < J->max_out; q++)
                    J->idx_out[t * J->max_out + q] = q < c ? (int32_t)w.nb[q] : -1;
                continue;
            }
            double v, var, lag;
            int hl;
            int32_t st;
            if (J->block)
                st = krig_predict_block(J->K, tg, J->offsets, J->n_off, &w, &v, &var, &lag, &hl);
            else if (J->tdrift)   /* predict_with_drift (kriging.rs:505-522) */
                st = J->K->cfg.method != RC_EXTERNAL_DRIFT
                         ? RC_INVALID_PARAMETER
                         : krig_predict(J->K, tg, J->tdrift + t * J->K->nv, &w, &v, &var, &lag, &hl);
            else                  /* predict (kriging.rs:494-501) */
                st = J->K->cfg.method == RC_EXTERNAL_DRIFT
                         ? RC_INVALID_PARAMETER
                         : krig_predict(J->K, tg, NULL, &w, &v, &var, &lag, &hl);
            if (st != RC_OK) { v = NAN; var = NAN; lag = NAN; } /* NAN_ESTIMATE, kriging.rs:130-134 */
            J->value[t] = v;
            J->variance[t] = var;
            if (J->lagrange) J->lagrange[t] = hl ? lag : NAN;
            if (J->status) J->status[t] = st;
        }
    }
    kws_free(&w);
    return NULL;
}

static void run_kjob(kjob *J, int32_t nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    for (int i = 1; i < nthreads; i++) pthread_create(&th[i], NULL, kworker, J);
    kworker(J);
    for (int i = 1; i < nthreads; i++) pthread_join(th[i], NULL);
}

int32_t rc_kriging_predict_many(int32_t D, const double *coords, const double *values, int64_t n,
                                const rc_model *model, const rc_kconfig *cfg, const double *targets,
                                int64_t T, int32_t nthreads, double *value, double *variance,
                                double *lagrange, int32_t *status, int64_t *dup_pair) {
    krig K;
    int32_t st = krig_build(&K, D, coords, values, n, model, cfg, dup_pair);
    if (st != RC_OK) return st;
    kjob J;
    memset(&J, 0, sizeof(J));
    J.K = &K; J.targets = targets; J.T = T;
    atomic_init(&J.next, 0);
    J.value = value; J.variance = variance; J.lagrange = lagrange; J.status = status;
    run_kjob(&J, nthreads);
    krig_free(&K);
    return RC_OK;
}

/* with_measurement_error (kriging.rs:181-218) / with_external_drift
 * (223-260) + predict_many (648-652), predict_many_with_drift (656-672) or
 * predict_block per target (678-777, as predict_block_grid maps it). */
int32_t rc_kriging_predict_many_ex(int32_t D, const double *coords, const double *values,
                                   int64_t n, const rc_model *model, const rc_kconfig *cfg,
                                   const double *errors, int32_t n_vars, const double *drift_data,
                                   const double *targets, int64_t T, const double *tdrift,
                                   const double *offsets, int64_t n_off, int32_t block,
                                   int32_t nthreads, double *value, double *variance,
                                   double *lagrange, int32_t *status, int64_t *dup_pair) {
    if (errors) {
        if (cfg->method == RC_EXTERNAL_DRIFT) return RC_INVALID_PARAMETER;
        if (has_power(model)) return RC_INVALID_PARAMETER;
        for (int64_t i = 0; i < n; i++)
            if (!(isfinite(errors[i]) && errors[i] >= 0.0)) return RC_INVALID_PARAMETER;
    }
    if (drift_data) {
        if (cfg->method != RC_EXTERNAL_DRIFT) return RC_INVALID_PARAMETER;
        if (n_vars == 0) return RC_INVALID_PARAMETER;
        if (n_vars != cfg->degree || n_vars > 8) return RC_DIMENSION_MISMATCH;
        for (int64_t i = 0; i < n * n_vars; i++)
            if (!isfinite(drift_data[i])) return RC_INVALID_PARAMETER;
    }
    krig K;
    int32_t st = krig_build(&K, D, coords, values, n, model, cfg, dup_pair);
    if (st != RC_OK) return st;
    K.err = errors;
    if (drift_data) {
        K.nv = n_vars;
        K.drift = drift_data;
        double nn = (double)n;
        for (int k = 0; k < n_vars; k++) {   /* kriging.rs:349-364 */
            double sum = -0.0;
            for (int64_t i = 0; i < n; i++) sum += drift_data[i * n_vars + k];
            double mean = sum / nn;
            double ss = -0.0;
            for (int64_t i = 0; i < n; i++) {
                double d = drift_data[i * n_vars + k] - mean;
                ss += d * d;
            }
            K.dmean[k] = mean;
            K.dscale[k] = fmax(sqrt(ss / nn), 2.2250738585072014e-308);
        }
    }
    kjob J;
    memset(&J, 0, sizeof(J));
    J.K = &K; J.targets = targets; J.T = T;
    J.tdrift = tdrift; J.offsets = offsets; J.n_off = n_off; J.block = block;
    atomic_init(&J.next, 0);
    J.value = value; J.variance = variance; J.lagrange = lagrange; J.status = status;
    run_kjob(&J, nthreads);
    krig_free(&K);
    return RC_OK;
}

int32_t rc_kriging_neighbors(int32_t D, const double *coords, const double *values, int64_t n,
                             const rc_model *model, const rc_kconfig *cfg, const double *targets,
                             int64_t T, int64_t max_out, int32_t *idx_out, int32_t *count_out,
                             int32_t nthreads) {
    krig K;
    int64_t dup[2];
    int32_t st = krig_build(&K, D, coords, values, n, model, cfg, dup);
    if (st != RC_OK) return st;
    kjob J;
    memset(&J, 0, sizeof(J));
    J.K = &K; J.targets = targets; J.T = T;
    atomic_init(&J.next, 0);
    J.max_out = max_out; J.idx_out = idx_out; J.count_out = count_out;
    run_kjob(&J, nthreads);
    krig_free(&K);
    return RC_OK;
}

/* ------------------------------------------------------------------ */
/* vecchia.rs                                                          */
/* ------------------------------------------------------------------ */
static double dist2v(const double *a, const double *b, int D) { /* vecchia.rs:95-102 */
    double s = 0.0;
    for (int d = 0; d < D; d++) {
        double v = a[d] - b[d];
        s += v * v;
    }
    return s;
}

static void centroid_first(int32_t D, const double *coords, int64_t n, int64_t *first) { /* vecchia.rs:150-160 */
    double c[3] = {0.0, 0.0, 0.0};
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < D; d++) c[d] += coords[i * D + d] / (double)n;
    int64_t best = 0;
    double bd = dist2v(coords, c, D);
    for (int64_t i = 1; i < n; i++) {
        double dd = dist2v(coords + i * D, c, D);
        if (total_key(dd) < total_key(bd)) { bd = dd; best = i; } /* min_by keeps the first */
    }
    *first = best;
}

/* MaxminKey heap: max by (d2 total_cmp, idx) (vecchia.rs:104-132). */
typedef struct { double d2; int64_t idx; } mmkey;
static inline int mm_gt(mmkey a, mmkey b) {
    int64_t ka = total_key(a.d2), kb = total_key(b.d2);
    return ka > kb || (ka == kb && a.idx > b.idx);
}
typedef struct { mmkey *v; int64_t len; } mmheap;
static void mm_push(mmheap *h, mmkey e) {
    int64_t i = h->len++;
    h->v[i] = e;
    while (i > 0) {
        int64_t p = (i - 1) / 2;
        if (mm_gt(h->v[i], h->v[p])) { mmkey t = h->v[i]; h->v[i] = h->v[p]; h->v[p] = t; i = p; }
        else break;
    }
}
static mmkey mm_pop(mmheap *h) {
    mmkey top = h->v[0];
    h->v[0] = h->v[--h->len];
    int64_t i = 0;
    for (;;) {
        int64_t l = 2 * i + 1, r = l + 1, g = i;
        if (l < h->len && mm_gt(h->v[l], h->v[g])) g = l;
        if (r < h->len && mm_gt(h->v[r], h->v[g])) g = r;
        if (g == i) break;
        mmkey t = h->v[i]; h->v[i] = h->v[g]; h->v[g] = t; i = g;
    }
    return top;
}

static void bbox_of(int32_t D, const double *coords, int64_t n, double *mn, double *mx) {
    for (int d = 0; d < D; d++) { mn[d] = coords[d]; mx[d] = coords[d]; }
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < D; d++) {
            mn[d] = fmin(mn[d], coords[i * D + d]);
            mx[d] = fmax(mx[d], coords[i * D + d]);
        }
}

int32_t rc_maxmin_order(int32_t D, const double *coords, int64_t n, int64_t *order) { /* vecchia.rs:144-215 */
    if (n == 0) return RC_OK;
    int64_t first;
    centroid_first(D, coords, n, &first);
    double mn[3], mx[3];
    bbox_of(D, coords, n, mn, mx);
    bgrid g;
    bgrid_new(&g, D, mn, mx, n);
    double *sel = (double *)malloc(sizeof(double) * (size_t)(n * D));
    int64_t nsel = 0, no = 0;
    order[no++] = first;
    bgrid_insert(&g, coords + first * D);
    memcpy(sel, coords + first * D, sizeof(double) * D);
    nsel = 1;
    mmheap h;
    h.v = (mmkey *)malloc(sizeof(mmkey) * (size_t)(n + 1));
    h.len = 0;
    for (int64_t i = 0; i < n; i++)
        if (i != first) { mmkey e = {dist2v(coords + i * D, coords + first * D, D), i}; mm_push(&h, e); }
    mheap kh;
    kh.v = (hent *)malloc(sizeof(hent) * 2);
    kh.cap = 2;
    uint32_t nbq[2];
    while (h.len > 0) {
        mmkey top = mm_pop(&h);
        int64_t i = top.idx;
        double current;
        if (nsel <= 48) {
            current = INFINITY;
            for (int64_t s = 0; s < nsel; s++) current = fmin(current, dist2v(coords + i * D, sel + s * D, D));
        } else {
            bgrid_knn(&g, coords + i * D, 1, INFINITY, &kh, nbq);
            current = dist2v(coords + i * D, sel + (int64_t)nbq[0] * D, D);
        }
        mmkey refreshed = {current, i};
        if (h.len == 0 || !mm_gt(h.v[0], refreshed)) { /* refreshed >= next */
            order[no++] = i;
            bgrid_insert(&g, coords + i * D);
            memcpy(sel + nsel * D, coords + i * D, sizeof(double) * D);
            nsel++;
        } else mm_push(&h, refreshed);
    }
    free(kh.v);
    free(h.v);
    free(sel);
    bgrid_free(&g);
    return RC_OK;
}

/* maxmin_order_brute (vecchia.rs:2244-2273): the O(n^2) exact traversal. */
int32_t rc_maxmin_order_brute(int32_t D, const double *coords, int64_t n, int64_t *order) {
    if (n == 0) return RC_OK;
    int64_t first;
    centroid_first(D, coords, n, &first);
    char *chosen = (char *)calloc((size_t)n, 1);
    double *md = (double *)malloc(sizeof(double) * (size_t)n);
    order[0] = first;
    chosen[first] = 1;
    for (int64_t i = 0; i < n; i++) md[i] = dist2v(coords + i * D, coords + first * D, D);
    for (int64_t s = 1; s < n; s++) {
        int64_t next = -1;
        for (int64_t i = 0; i < n; i++) {
            if (chosen[i]) continue;
            if (next < 0 || total_key(md[i]) >= total_key(md[next])) next = i; /* max_by keeps the last */
        }
        order[s] = next;
        chosen[next] = 1;
        for (int64_t i = 0; i < n; i++)
            if (!chosen[i]) md[i] = fmin(md[i], dist2v(coords + i * D, coords + next * D, D));
    }
    free(chosen);
    free(md);
    return RC_OK;
}

/* vecchia_plan (vecchia.rs:232-282): nb rows are in ordered position. */
int32_t rc_vecchia_plan(int32_t D, const double *coords, int64_t n, int64_t m, const int64_t *order_in,
                        int64_t *order_out, int32_t *nb, int32_t *nb_count) {
    if (n < 2) return RC_INSUFFICIENT_DATA;
    if (m <= 0) return RC_INVALID_PARAMETER;
    int64_t dup[2];
    if (rc_duplicate_pair(D, coords, n, dup)) return RC_DUPLICATE_POINTS;
    if (order_in) memcpy(order_out, order_in, sizeof(int64_t) * (size_t)n);
    else rc_maxmin_order(D, coords, n, order_out);
    double mn[3], mx[3];
    bbox_of(D, coords, n, mn, mx);
    bgrid g;
    bgrid_new(&g, D, mn, mx, n);
    mheap h;
    h.v = (hent *)malloc(sizeof(hent) * (size_t)(m + 1));
    h.cap = m + 1;
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(m + 1));
    for (int64_t k = 0; k < n; k++) {
        int64_t i = order_out[k];
        int64_t c = bgrid_knn(&g, coords + i * D, m, INFINITY, &h, tmp);
        nb_count[k] = (int32_t)c;
        for (int64_t q = 0; q < m; q++) nb[k * m + q] = q < c ? (int32_t)order_out[tmp[q]] : -1;
        bgrid_insert(&g, coords + i * D);
    }
    free(tmp);
    free(h.v);
    bgrid_free(&g);
    return RC_OK;
}

/* loglik_with_plan (vecchia.rs:285-336). */
int32_t rc_vecchia_loglik_plan(int32_t D, const double *coords, const double *values, int64_t n,
                               const rc_model *model, const int64_t *order, const int32_t *nb,
                               const int32_t *nb_count, int64_t m, double *out) {
    if (has_power(model)) return RC_INVALID_PARAMETER;         /* vecchia.rs:44-49 */
    for (int i = 0; i < model->n_structures; i++)
        if (model->s[i].kind == RC_CIRCULAR && D > 2) return RC_INVALID_PARAMETER;
    double mean = -0.0;
    for (int64_t i = 0; i < n; i++) mean += values[i];
    mean = mean / (double)n;
    double sill = rc_total_sill(model);
    double *kss = (double *)malloc(sizeof(double) * (size_t)(m * m));
    double *ki = (double *)malloc(sizeof(double) * (size_t)m);
    double *w = (double *)malloc(sizeof(double) * (size_t)m);
    double ll = 0.0;
    double ln2pi = log(2.0 * M_PI);
    int32_t st = RC_OK;
    for (int64_t k = 0; k < n; k++) {
        int64_t i = order[k];
        int64_t s = nb_count[k];
        double mu, var;
        if (s == 0) { mu = 0.0; var = sill; }
        else {
            const int32_t *row = nb + k * m;
            for (int64_t a = 0; a < s; a++) {
                double dh[3];
                for (int d = 0; d < D; d++) dh[d] = coords[i * D + d] - coords[(int64_t)row[a] * D + d];
                ki[a] = sill - rc_gamma_dh(model, D, dh);
                for (int64_t b = a; b < s; b++) {
                    for (int d = 0; d < D; d++)
                        dh[d] = coords[(int64_t)row[a] * D + d] - coords[(int64_t)row[b] * D + d];
                    double c = sill - rc_gamma_dh(model, D, dh);
                    kss[a * s + b] = c;
                    kss[b * s + a] = c;
                }
            }
            memcpy(w, ki, sizeof(double) * (size_t)s);
            st = rc_cholesky_solve(kss, (int32_t)s, w);
            if (st != RC_OK) break;
            mu = -0.0;
            double red = -0.0;
            for (int64_t a = 0; a < s; a++) mu += w[a] * (values[row[a]] - mean);
            for (int64_t a = 0; a < s; a++) red += w[a] * ki[a];
            var = fmax(sill - red, 1e-12);
        }
        double r = (values[i] - mean) - mu;
        ll += -0.5 * (ln2pi + log(var) + r * r / var);
    }
    free(kss); free(ki); free(w);
    if (st != RC_OK) return st;
    *out = ll;
    return RC_OK;
}

/* g_and_dg (vecchia.rs:353-385): g(d) and dg/dd for the kinds with a
 * closed-form range derivative; returns 0 for the others. */
static int g_and_dg(int32_t kind, double d, double *g, double *dg) {
    switch (kind) {
    case RC_SPHERICAL:
        if (d >= 1.0) { *g = 1.0; *dg = 0.0; }
        else { *g = 1.5 * d - 0.5 * d * d * d; *dg = 1.5 - 1.5 * d * d; }
        return 1;
    case RC_EXPONENTIAL: { double e = exp(-d); *g = 1.0 - e; *dg = e; return 1; }
    case RC_GAUSSIAN: { double e = exp(-(d * d)); *g = 1.0 - e; *dg = 2.0 * d * e; return 1; }
    case RC_MATERN15: {
        double q = sqrt(3.0) * d, e = exp(-q);
        *g = 1.0 - (1.0 + q) * e;
        *dg = 3.0 * d * e;
        return 1;
    }
    case RC_MATERN25: {
        double q = sqrt(5.0) * d, e = exp(-q);
        *g = 1.0 - (1.0 + q + q * q / 3.0) * e;
        *dg = 5.0 * d * (1.0 + q) / 3.0 * e;
        return 1;
    }
    default: return 0;
    }
}

static double vdist2(const double *a, const double *b, int D) {   /* vecchia.rs:95-102 */
    double s = 0.0;
    for (int d = 0; d < D; d++) { double v = a[d] - b[d]; s += v * v; }
    return s;
}

/* cholesky_factor_in_place (linalg.rs:123-160) and the two triangular
 * solves (linalg.rs:162-203), split so one factor serves several
 * right-hand sides. */
static int32_t chol_factor(double *a, int32_t n) {
    double scale = 0.0;
    for (int i = 0; i < n; i++) scale = fmax(scale, a[i * n + i]);
    scale = fmax(scale, 2.2250738585072014e-308);
    for (int k = 0; k < n; k++) {
        double d = a[k * n + k];
        for (int j = 0; j < k; j++) d -= a[k * n + j] * a[k * n + j];
        if (!(d > 1e-12 * scale)) return RC_SINGULAR_SYSTEM;
        d = sqrt(d);
        a[k * n + k] = d;
        for (int i = k + 1; i < n; i++) {
            double s = a[i * n + k];
            for (int j = 0; j < k; j++) s -= a[i * n + j] * a[k * n + j];
            a[i * n + k] = s / d;
        }
    }
    return RC_OK;
}
static void chol_solve(const double *a, int32_t n, double *b) {
    for (int i = 0; i < n; i++) {
        double s = b[i];
        for (int j = 0; j < i; j++) s -= a[i * n + j] * b[j];
        b[i] = s / a[i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int j = i + 1; j < n; j++) s -= a[j * n + i] * b[j];
        b[i] = s / a[i * n + i];
    }
}

/* loglik_grad_inner (vecchia.rs:428-566): log-likelihood and its gradient
 * in (nugget, psill, range) for one isotropic structure of a kind with a
 * closed-form range derivative.  Dense per-point matrices K_SS, dK/dpsill,
 * dK/drange; one factorization reused for w = K^-1 k_i and the three
 * derivative solves; the variance term's derivative is zeroed where the
 * variance is clamped at 1e-12. */
int32_t rc_vecchia_loglik_grad_plan(int32_t D, const double *coords, const double *values,
                                    int64_t n, int32_t kind, double nugget, double psill,
                                    double range, const int64_t *order, const int32_t *nb,
                                    const int32_t *nb_count, int64_t m, double *ll_out,
                                    double *grad_out) {
    double g0, dg0;
    if (!g_and_dg(kind, 0.5, &g0, &dg0)) return RC_INVALID_PARAMETER;
    double mean = -0.0;
    for (int64_t i = 0; i < n; i++) mean += values[i];
    mean = mean / (double)n;
    const double sill = nugget + psill;
    const size_t mm = (size_t)(m * m);
    double *kss = (double *)malloc(sizeof(double) * mm);
    double *dkp = (double *)malloc(sizeof(double) * mm);
    double *dkr = (double *)malloc(sizeof(double) * mm);
    double *vec = (double *)malloc(sizeof(double) * (size_t)(8 * m));
    double *ki = vec, *dkip = vec + m, *dkir = vec + 2 * m, *w = vec + 3 * m;
    double *rhs = vec + 4 * m, *dw = vec + 5 * m, *zn = vec + 6 * m, *zeros = vec + 7 * m;
    for (int64_t a = 0; a < m; a++) zeros[a] = 0.0;
    const double ln2pi = log(2.0 * M_PI);
    double ll = 0.0, grad[3] = {0.0, 0.0, 0.0};
    int32_t st = RC_OK;
    for (int64_t k = 0; k < n && st == RC_OK; k++) {
        const int64_t i = order[k];
        const int64_t s = nb_count[k];
        const double zi = values[i] - mean;
        if (s == 0) {
            const double var = sill, r = zi;
            ll += -0.5 * (ln2pi + log(var) + r * r / var);
            const double common = -0.5 / var + 0.5 * r * r / (var * var);
            grad[0] += common;
            grad[1] += common;
            continue;
        }
        const int32_t *row = nb + k * m;
        for (int64_t a = 0; a < s; a++) {
            zn[a] = values[row[a]] - mean;
            double g, dg;
            const double h_ia = sqrt(vdist2(coords + i * D, coords + (int64_t)row[a] * D, D));
            g_and_dg(kind, h_ia / range, &g, &dg);
            ki[a] = sill - (nugget + psill * g);
            dkip[a] = 1.0 - g;
            dkir[a] = psill * dg * h_ia / (range * range);
            kss[a * s + a] = sill;
            dkp[a * s + a] = 1.0;
            dkr[a * s + a] = 0.0;
            for (int64_t b = a + 1; b < s; b++) {
                const double h = sqrt(vdist2(coords + (int64_t)row[a] * D, coords + (int64_t)row[b] * D, D));
                g_and_dg(kind, h / range, &g, &dg);
                const double c = sill - (nugget + psill * g);
                kss[a * s + b] = kss[b * s + a] = c;
                dkp[a * s + b] = dkp[b * s + a] = 1.0 - g;
                const double dr = psill * dg * h / (range * range);
                dkr[a * s + b] = dkr[b * s + a] = dr;
            }
        }
        st = chol_factor(kss, (int32_t)s);
        if (st != RC_OK) break;
        memcpy(w, ki, sizeof(double) * (size_t)s);
        chol_solve(kss, (int32_t)s, w);
        double mu = 0.0, red = 0.0;
        for (int64_t a = 0; a < s; a++) mu += w[a] * zn[a];
        for (int64_t a = 0; a < s; a++) red += w[a] * ki[a];
        const double raw = sill - red;
        const double var = fmax(raw, 1e-12);
        const int clamped = raw <= 1e-12;
        const double r = zi - mu;
        ll += -0.5 * (ln2pi + log(var) + r * r / var);
        for (int p = 0; p < 3; p++) {
            const double dsill = p < 2 ? 1.0 : 0.0;
            const double *dk = p == 0 ? NULL : (p == 1 ? dkp : dkr);
            const double *dki = p == 0 ? zeros : (p == 1 ? dkip : dkir);
            for (int64_t a = 0; a < s; a++) {
                if (!dk) { rhs[a] = -w[a]; continue; }
                double y = 0.0;                                 /* matvec (vecchia.rs:388-395) */
                for (int64_t b = 0; b < s; b++) y += dk[a * s + b] * w[b];
                rhs[a] = dki[a] - y;
            }
            memcpy(dw, rhs, sizeof(double) * (size_t)s);
            chol_solve(kss, (int32_t)s, dw);
            double dmu = 0.0, dr1 = 0.0, dr2 = 0.0;
            for (int64_t a = 0; a < s; a++) dmu += dw[a] * zn[a];
            for (int64_t a = 0; a < s; a++) dr1 += dw[a] * ki[a];
            for (int64_t a = 0; a < s; a++) dr2 += w[a] * dki[a];
            const double dvar = clamped ? 0.0 : dsill - (dr1 + dr2);
            grad[p] += -0.5 * dvar / var + r * dmu / var + 0.5 * r * r / (var * var) * dvar;
        }
    }
    free(kss); free(dkp); free(dkr); free(vec);
    if (st != RC_OK) return st;
    *ll_out = ll;
    for (int p = 0; p < 3; p++) grad_out[p] = grad[p];
    return RC_OK;
}

/* exact_loglik (vecchia.rs:1958-1974): dense MVN loglik through LU. */
int32_t rc_dense_mvn_loglik(int32_t D, const double *coords, const double *values, int64_t n,
                            const rc_model *model, double *out) {
    double mean = -0.0;
    for (int64_t i = 0; i < n; i++) mean += values[i];
    mean = mean / (double)n;
    double *z = (double *)malloc(sizeof(double) * (size_t)n);
    double *x = (double *)malloc(sizeof(double) * (size_t)n);
    double *C = (double *)malloc(sizeof(double) * (size_t)(n * n));
    double c0 = rc_total_sill(model);
    for (int64_t i = 0; i < n; i++) z[i] = values[i] - mean;
    for (int64_t a = 0; a < n; a++)
        for (int64_t b = 0; b < n; b++) {
            double dh[3];
            for (int d = 0; d < D; d++) dh[d] = coords[a * D + d] - coords[b * D + d];
            C[a * n + b] = c0 - rc_gamma_dh(model, D, dh);
        }
    memcpy(x, z, sizeof(double) * (size_t)n);
    double logdet;
    int32_t st = lu_solve_logdet(C, (int32_t)n, x, &logdet);
    if (st == RC_OK) {
        double q = -0.0;
        for (int64_t i = 0; i < n; i++) q += z[i] * x[i];
        *out = -0.5 * ((double)n * log(2.0 * M_PI) + logdet + q);
    }
    free(z); free(x); free(C);
    return st;
}

/* ---------------------------------------------------------------- optim.rs */
/* Objective callback: f(x) and, when g is non-NULL, the gradient. */
typedef struct {
    int32_t D;
    const double *coords, *values;
    int64_t n, m;
    const int64_t *order;
    const int32_t *nb, *nb_count;
    int32_t kind;
    double param, range_max;
    int64_t evals;
} mle_ctx;

static double mle_reg(const mle_ctx *c, double range) {   /* vecchia.rs:988-996 */
    if (range > c->range_max) {
        double t = (range - c->range_max) / c->range_max;
        return t * t;
    }
    return 0.0;
}

static double mle_f(mle_ctx *c, const double *x) {        /* vecchia.rs:998-1016 */
    rc_model md;
    memset(&md, 0, sizeof(md));
    md.nugget = x[0] * x[0];
    md.n_structures = 1;
    md.s[0].kind = c->kind;
    md.s[0].param = c->param;
    md.s[0].sill = exp(x[1]);
    md.s[0].range = exp(x[2]);
    double ll;
    c->evals++;
    int32_t st = rc_vecchia_loglik_plan(c->D, c->coords, c->values, c->n, &md, c->order, c->nb,
                                        c->nb_count, c->m, &ll);
    if (st == RC_OK && isfinite(ll)) return -ll + 1e6 * mle_reg(c, md.s[0].range);
    return 1e12;
}

static void mle_g(mle_ctx *c, const double *x, double *gx) { /* vecchia.rs:1030-1050 */
    const double nugget = x[0] * x[0], psill = exp(x[1]), range = exp(x[2]);
    double ll, g[3];
    c->evals++;
    int32_t st = rc_vecchia_loglik_grad_plan(c->D, c->coords, c->values, c->n, c->kind, nugget,
                                             psill, range, c->order, c->nb, c->nb_count, c->m,
                                             &ll, g);
    if (st != RC_OK) { gx[0] = gx[1] = gx[2] = 0.0; return; }
    gx[0] = -g[0] * 2.0 * x[0];
    gx[1] = -g[1] * psill;
    gx[2] = -g[2] * range;
    if (range > c->range_max) {
        double dreg = 2.0 * (range - c->range_max) / (c->range_max * c->range_max);
        gx[2] += 1e6 * dreg * range;
    }
}

static int f64_total_lt(double a, double b) { return total_key(a) < total_key(b); }

/* nelder_mead (optim.rs:6-99) for n = 3 parameters. */
static double nelder_mead3(mle_ctx *c, const double *x0, double step, int max_iter, double *xout) {
    enum { N = 3 };
    double sx[N + 1][N], fv[N + 1];
    for (int j = 0; j < N; j++) sx[0][j] = x0[j];
    for (int i = 0; i < N; i++) {
        for (int j = 0; j < N; j++) sx[i + 1][j] = x0[j];
        sx[i + 1][i] += step;
    }
    for (int i = 0; i <= N; i++) fv[i] = mle_f(c, sx[i]);
    for (int it = 0; it < max_iter; it++) {
        int idx[N + 1];
        for (int i = 0; i <= N; i++) idx[i] = i;
        for (int i = 1; i <= N; i++)   /* stable insertion sort by total_cmp */
            for (int j = i; j > 0 && f64_total_lt(fv[idx[j]], fv[idx[j - 1]]); j--) {
                int t = idx[j]; idx[j] = idx[j - 1]; idx[j - 1] = t;
            }
        double s2[N + 1][N], f2[N + 1];
        for (int i = 0; i <= N; i++) {
            memcpy(s2[i], sx[idx[i]], sizeof(s2[i]));
            f2[i] = fv[idx[i]];
        }
        memcpy(sx, s2, sizeof(sx));
        memcpy(fv, f2, sizeof(fv));
        if (fabs(fv[N] - fv[0]) <= 1e-12 * fmax(fabs(fv[0]), 1e-12)) break;
        double cen[N] = {0.0, 0.0, 0.0};
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++) cen[j] += sx[i][j] / (double)N;
        double xr[N];
        for (int j = 0; j < N; j++) xr[j] = cen[j] + 1.0 * (cen[j] - sx[N][j]);
        double fr = mle_f(c, xr);
        if (fr < fv[0]) {
            double xe[N];
            for (int j = 0; j < N; j++) xe[j] = cen[j] + 2.0 * (cen[j] - sx[N][j]);
            double fe = mle_f(c, xe);
            if (fe < fr) { memcpy(sx[N], xe, sizeof(xe)); fv[N] = fe; }
            else { memcpy(sx[N], xr, sizeof(xr)); fv[N] = fr; }
        } else if (fr < fv[N - 1]) {
            memcpy(sx[N], xr, sizeof(xr));
            fv[N] = fr;
        } else {
            double xc[N];
            if (fr < fv[N])
                for (int j = 0; j < N; j++) xc[j] = cen[j] + 0.5 * (xr[j] - cen[j]);
            else
                for (int j = 0; j < N; j++) xc[j] = cen[j] - 0.5 * (cen[j] - sx[N][j]);
            double fc = mle_f(c, xc);
            if (fc < fmin(fr, fv[N])) {
                memcpy(sx[N], xc, sizeof(xc));
                fv[N] = fc;
            } else {
                for (int i = 1; i <= N; i++) {
                    for (int j = 0; j < N; j++) sx[i][j] = sx[0][j] + 0.5 * (sx[i][j] - sx[0][j]);
                    fv[i] = mle_f(c, sx[i]);
                }
            }
        }
    }
    int best = 0;
    for (int i = 1; i <= N; i++)
        if (fv[i] < fv[best]) best = i;
    memcpy(xout, sx[best], sizeof(double) * N);
    return fv[best];
}

static double dot3(const double *a, const double *b) {   /* optim.rs:125-127 */
    double s = 0.0;
    for (int i = 0; i < 3; i++) s += a[i] * b[i];
    return s;
}

static void mat3_mul(const double *a, const double *b, double *out) { /* optim.rs:153-165 */
    double c[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s += a[i * 3 + k] * b[k * 3 + j];
            c[i * 3 + j] = s;
        }
    memcpy(out, c, sizeof(c));
}

static void bfgs_update3(double *h, const double *s, const double *y, double sy) { /* optim.rs:167-187 */
    const double rho = 1.0 / sy;
    double left[9], right[9], t[9];
    for (int i = 0; i < 9; i++) left[i] = right[i] = (i % 4 == 0) ? 1.0 : 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            left[i * 3 + j] -= rho * (s[i] * y[j]);
            right[i * 3 + j] -= rho * (y[i] * s[j]);
        }
    mat3_mul(left, h, t);
    mat3_mul(t, right, h);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) h[i * 3 + j] += rho * (s[i] * s[j]);
}

/* bfgs (optim.rs:206-269) for n = 3 parameters. */
static double bfgs3(mle_ctx *c, const double *x0, int max_iter, double *xout) {
    double x[3], h[9], g[3], bx[3];
    memcpy(x, x0, sizeof(x));
    for (int i = 0; i < 9; i++) h[i] = (i % 4 == 0) ? 1.0 : 0.0;
    mle_g(c, x, g);
    double fx = mle_f(c, x);
    memcpy(bx, x, sizeof(x));
    double bf = fx;
    for (int it = 0; it < max_iter; it++) {
        if (sqrt(dot3(g, g)) < 1e-10) break;
        double p[3];
        for (int i = 0; i < 3; i++) {
            double s = 0.0;
            for (int j = 0; j < 3; j++) s += h[i * 3 + j] * g[j];
            p[i] = -s;
        }
        double slope = dot3(g, p);
        if (slope >= 0.0) {
            for (int i = 0; i < 9; i++) h[i] = (i % 4 == 0) ? 1.0 : 0.0;
            for (int i = 0; i < 3; i++) p[i] = -g[i];
            slope = dot3(g, p);
        }
        double alpha = 1.0, xn[3], fn;
        for (;;) {
            for (int i = 0; i < 3; i++) xn[i] = x[i] + alpha * p[i];
            fn = mle_f(c, xn);
            if (fn <= fx + 1e-4 * alpha * slope || alpha < 1e-12) break;
            alpha *= 0.5;
        }
        double gn[3], s[3], y[3];
        mle_g(c, xn, gn);
        for (int i = 0; i < 3; i++) { s[i] = xn[i] - x[i]; y[i] = gn[i] - g[i]; }
        double sy = dot3(s, y);
        if (sy > 1e-10) bfgs_update3(h, s, y, sy);
        int stalled = fabs(fx - fn) <= 1e-13 * (1.0 + fabs(fx)) &&
                      sqrt(dot3(s, s)) <= 1e-10 * (1.0 + sqrt(dot3(x, x)));
        memcpy(x, xn, sizeof(x));
        memcpy(g, gn, sizeof(g));
        fx = fn;
        if (fx < bf) { bf = fx; memcpy(bx, x, sizeof(x)); }
        if (stalled) break;
    }
    memcpy(xout, bx, sizeof(bx));
    return bf;
}

/* mle_fit_with_plan (vecchia.rs:952-1081) with group_size 1, after the kind
 * checks of vecchia_mle (vecchia.rs:921-932).  fit = (nugget, psill, range). */
int32_t rc_vecchia_mle_plan(int32_t D, const double *coords, const double *values, int64_t n,
                            int32_t kind, double param, const int64_t *order, const int32_t *nb,
                            const int32_t *nb_count, int64_t m, double *fit, double *loglik,
                            int64_t *evals) {
    if (kind == RC_POWER) return RC_INVALID_PARAMETER;
    if (kind == RC_CIRCULAR && D > 2) return RC_INVALID_PARAMETER;
    double mean = -0.0;                                    /* PointSet::mean: sum / n */
    for (int64_t i = 0; i < n; i++) mean += values[i];
    mean = mean / (double)n;
    double ss = -0.0;
    for (int64_t i = 0; i < n; i++) ss += (values[i] - mean) * (values[i] - mean);
    const double var0 = fmax(ss / (double)n, 1e-12);
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < D; d++) {
            lo[d] = fmin(lo[d], coords[i * D + d]);
            hi[d] = fmax(hi[d], coords[i * D + d]);
        }
    double e2 = -0.0;
    for (int d = 0; d < D; d++) e2 += (hi[d] - lo[d]) * (hi[d] - lo[d]);
    const double extent = sqrt(e2);
    const double range0 = fmax(extent / 3.0, 1e-9);
    mle_ctx c = {D, coords, values, n, m, order, nb, nb_count, kind, param, 10.0 * extent, 0};
    const double facs[3] = {0.3, 1.0, 3.0};
    double g0, dg0;
    const int use_bfgs = g_and_dg(kind, 0.5, &g0, &dg0);
    double xb[3] = {0, 0, 0}, fb = 0.0;
    int have = 0;
    for (int s = 0; s < 3; s++) {   /* bfgs_multistart / nelder_mead_multistart (optim.rs:106-123, 275-288) */
        double x0[3] = {sqrt(0.1 * var0), log(0.9 * var0), log(range0) + log(facs[s])};
        double xs[3], fs;
        fs = use_bfgs ? bfgs3(&c, x0, 200, xs) : nelder_mead3(&c, x0, 0.3, 2000, xs);
        if (!have || fs < fb) { memcpy(xb, xs, sizeof(xb)); fb = fs; have = 1; }
    }
    if (evals) *evals = c.evals;
    if (fb >= 1e12) return RC_SINGULAR_SYSTEM;
    fit[0] = xb[0] * xb[0];
    fit[1] = exp(xb[1]);
    fit[2] = exp(xb[2]);
    *loglik = -fb;
    return RC_OK;
}
