/*
 * gs_synth.c — seeded synthetic inputs for benchmarks and tests.
 *
 * The generator is xoshiro256++ seeded by splitmix64 with a 53-bit uniform
 * and a Box-Muller normal that returns r*cos first and caches r*sin, the
 * same stream definition as geostat-core's Rng (rng.rs:8-75), so data drawn
 * here equals data drawn by the reference's own bench recipe
 * (benches/engine.rs:12-23: x = U*L, y = U*L, v = sin(x/150) + cos(y/200) +
 * s*N).  Host-only C; no GPU code and nothing from the test oracle.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { uint64_t s[4]; int spare_ok; double spare; } sy_rng;

static uint64_t sy_splitmix(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static void sy_seed(sy_rng *r, uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) r->s[i] = sy_splitmix(&x);
    r->spare_ok = 0;
}
static inline uint64_t sy_rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
static uint64_t sy_next(sy_rng *r) {
    uint64_t *s = r->s;
    const uint64_t out = sy_rotl(s[0] + s[3], 23) + s[0];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = sy_rotl(s[3], 45);
    return out;
}
static double sy_uniform(sy_rng *r) { return (double)(sy_next(r) >> 11) * 0x1.0p-53; }
static double sy_normal(sy_rng *r) {
    if (r->spare_ok) { r->spare_ok = 0; return r->spare; }
    for (;;) {
        double u1 = sy_uniform(r);
        if (!(u1 > 0.0)) continue;
        double u2 = sy_uniform(r);
        double rad = sqrt(-2.0 * log(u1));
        double th = 2.0 * M_PI * u2;
        r->spare = rad * sin(th);
        r->spare_ok = 1;
        return rad * cos(th);
    }
}

/* First `count` raw draws of the stream (pins the generator in tests). */
void gs_synth_raw(uint64_t seed, int64_t count, uint64_t *out) {
    sy_rng r;
    sy_seed(&r, seed);
    for (int64_t i = 0; i < count; i++) out[i] = sy_next(&r);
}

void gs_synth_uniform(uint64_t seed, int64_t count, double *out) {
    sy_rng r;
    sy_seed(&r, seed);
    for (int64_t i = 0; i < count; i++) out[i] = sy_uniform(&r);
}

/*
 * recipe 0: 2-D, per point x = U*ext0, y = U*ext1,
 *           v = sin(x/150) + cos(y/200) + noise*N
 * recipe 1: 3-D, per point x, y, z = U*ext, then
 *           v = 0.001*z + sin(x/150) + cos(y/200) + noise*N
 * recipe 2: coordinates only (U*ext per axis), values 0
 */
int gs_synth_points(uint64_t seed, int64_t n, int dim, const double *ext, int recipe,
                    double noise, double *coords, double *values) {
    sy_rng r;
    sy_seed(&r, seed);
    for (int64_t i = 0; i < n; i++) {
        double c[3] = {0, 0, 0};
        for (int d = 0; d < dim; d++) c[d] = sy_uniform(&r) * ext[d];
        for (int d = 0; d < dim; d++) coords[i * dim + d] = c[d];
        double v = 0.0;
        if (recipe == 0) v = sin(c[0] / 150.0) + cos(c[1] / 200.0) + noise * sy_normal(&r);
        else if (recipe == 1)
            v = 0.001 * c[2] + sin(c[0] / 150.0) + cos(c[1] / 200.0) + noise * sy_normal(&r);
        values[i] = v;
    }
    return 0;
}

/* ---- seeded Vecchia-approximate simulation of a Matern-2.5 field ------ */

static double mat25_cov(double h, double nugget, double sill, double range) {
    if (h == 0.0) return nugget + sill;
    double q = sqrt(5.0) * h / range;
    return sill * (1.0 + q + q * q / 3.0) * exp(-q);
}

typedef struct {
    int dim;
    const double *coords;
    const int64_t *order;
    const int32_t *nb, *nb_count;
    int64_t n, m;
    double nugget, sill, range;
    double *w;      /* [n][m] conditional weights */
    double *sd;     /* [n] conditional standard deviations */
    int64_t next;
    pthread_mutex_t mu;
} simjob;

static void *sim_worker(void *arg) {
    simjob *J = (simjob *)arg;
    const int64_t m = J->m;
    const int dim = J->dim;
    double *A = (double *)malloc(sizeof(double) * (size_t)(m * m));
    double *b = (double *)malloc(sizeof(double) * (size_t)m);
    for (;;) {
        pthread_mutex_lock(&J->mu);
        int64_t s = J->next;
        J->next += 512;
        pthread_mutex_unlock(&J->mu);
        if (s >= J->n) break;
        int64_t e = s + 512 < J->n ? s + 512 : J->n;
        for (int64_t k = s; k < e; k++) {
            int64_t i = J->order[k];
            int64_t c = J->nb_count[k];
            const int32_t *row = J->nb + k * m;
            double c0 = J->nugget + J->sill;
            for (int64_t a = 0; a < c; a++) {
                double h2 = 0.0;
                for (int d = 0; d < dim; d++) {
                    double t = J->coords[i * dim + d] - J->coords[(int64_t)row[a] * dim + d];
                    h2 += t * t;
                }
                b[a] = mat25_cov(sqrt(h2), J->nugget, J->sill, J->range);
                for (int64_t q = 0; q <= a; q++) {
                    double g2 = 0.0;
                    for (int d = 0; d < dim; d++) {
                        double t = J->coords[(int64_t)row[a] * dim + d] - J->coords[(int64_t)row[q] * dim + d];
                        g2 += t * t;
                    }
                    A[a * m + q] = mat25_cov(sqrt(g2), J->nugget, J->sill, J->range);
                }
            }
            /* Cholesky + solve A w = b; var = c0 - b'w */
            int ok = 1;
            for (int64_t a = 0; a < c && ok; a++) {
                double dd = A[a * m + a];
                for (int64_t q = 0; q < a; q++) dd -= A[a * m + q] * A[a * m + q];
                if (!(dd > 0.0)) { ok = 0; break; }
                dd = sqrt(dd);
                A[a * m + a] = dd;
                for (int64_t r = a + 1; r < c; r++) {
                    double s2 = A[r * m + a];
                    for (int64_t q = 0; q < a; q++) s2 -= A[r * m + q] * A[a * m + q];
                    A[r * m + a] = s2 / dd;
                }
            }
            double *w = J->w + k * m;
            double var = c0;
            if (ok && c > 0) {
                double *y = w;
                for (int64_t a = 0; a < c; a++) {
                    double s2 = b[a];
                    for (int64_t q = 0; q < a; q++) s2 -= A[a * m + q] * y[q];
                    y[a] = s2 / A[a * m + a];
                }
                double red = 0.0;
                for (int64_t a = 0; a < c; a++) red += y[a] * y[a];
                for (int64_t a = c - 1; a >= 0; a--) {
                    double s2 = y[a];
                    for (int64_t q = a + 1; q < c; q++) s2 -= A[q * m + a] * w[q];
                    w[a] = s2 / A[a * m + a];
                }
                var = c0 - red;
            } else {
                for (int64_t a = 0; a < m; a++) w[a] = 0.0;
            }
            J->sd[k] = sqrt(var > 1e-12 ? var : 1e-12);
        }
    }
    free(A);
    free(b);
    return NULL;
}

/* Draws v[order[k]] = sum_j w_kj v[nb_kj] + sd_k * N sequentially over the
 * ordering, with weights and sds from the conditional Gaussian of each point
 * given its plan neighbours (zero mean, nugget + Matern-2.5). */
int gs_synth_vecchia_field(uint64_t seed, int64_t n, int dim, const double *coords,
                           const int64_t *order, const int32_t *nb, const int32_t *nb_count,
                           int64_t m, double nugget, double sill, double range, int nthreads,
                           double *values) {
    simjob J;
    memset(&J, 0, sizeof(J));
    J.dim = dim; J.coords = coords; J.order = order; J.nb = nb; J.nb_count = nb_count;
    J.n = n; J.m = m; J.nugget = nugget; J.sill = sill; J.range = range;
    J.w = (double *)malloc(sizeof(double) * (size_t)(n * m));
    J.sd = (double *)malloc(sizeof(double) * (size_t)n);
    pthread_mutex_init(&J.mu, NULL);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    pthread_t th[64];
    for (int i = 1; i < nthreads; i++) pthread_create(&th[i], NULL, sim_worker, &J);
    sim_worker(&J);
    for (int i = 1; i < nthreads; i++) pthread_join(th[i], NULL);
    sy_rng r;
    sy_seed(&r, seed);
    for (int64_t k = 0; k < n; k++) {
        int64_t i = order[k];
        double mu = 0.0;
        for (int64_t a = 0; a < nb_count[k]; a++) mu += J.w[k * m + a] * values[nb[k * m + a]];
        values[i] = mu + J.sd[k] * sy_normal(&r);
    }
    pthread_mutex_destroy(&J.mu);
    free(J.w);
    free(J.sd);
    return 0;
}
