// Vecchia maximum-likelihood fit (vecchia_mle / mle_fit_with_plan,
// vecchia.rs:921-1081) on top of the batched GPU likelihood.
//
// The reference runs its optimizer (optim.rs) once per start, one
// likelihood evaluation at a time.  The starts are independent, so here they
// advance in lockstep: every round collects the point each unfinished start
// needs next and evaluates all of them in one batched launch.  Each start
// follows exactly the reference's sequence of steps (BFGS with Armijo
// backtracking, optim.rs:206-288, for the kinds with a closed-form range
// derivative; Nelder-Mead, optim.rs:6-123, otherwise), so the per-start
// trajectories and the final pick (best objective, first start on ties) are
// the reference's, up to the rounding of the likelihood itself.  A BFGS
// evaluation returns the objective and its gradient together, so the
// gradient at an accepted line-search point needs no extra launch.
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/geostat_hip.h"
#include "gs_internal.hpp"

namespace {

constexpr int kN = 3;   // x = (sqrt nugget, ln psill, ln range)
using Vec = std::array<double, kN>;
using Mat = std::array<double, kN * kN>;

// Parameters of one single-structure model (mle_fit_with_plan objective).
gs_model model_of(const Vec &x, int32_t kind, double param) {
    gs_model m;
    std::memset(&m, 0, sizeof(m));
    m.nugget = x[0] * x[0];
    m.n_structures = 1;
    m.structures[0].kind = kind;
    m.structures[0].param = param;
    m.structures[0].sill = std::exp(x[1]);
    m.structures[0].range = std::exp(x[2]);
    m.structures[0].has_anis = 0;
    return m;
}

// A candidate whose parameters overflow or underflow (exp of a far-off
// x[1] or x[2]) has no likelihood: the objective's failure value.
bool model_ok(const gs_model &m) {
    const double s = m.structures[0].sill, r = m.structures[0].range;
    return std::isfinite(m.nugget) && std::isfinite(s) && s > 0.0 && std::isfinite(r) && r > 0.0;
}

// f64::total_cmp as an integer key
int64_t total_key(double x) {
    int64_t b;
    std::memcpy(&b, &x, sizeof(b));
    return b ^ (int64_t)((uint64_t)(b >> 63) >> 1);
}

struct Evaluator {
    gs_vplan *plan;
    const double *values;
    int32_t kind;
    double param;
    double range_max;
    int32_t err = GS_OK;

    double reg(double range) const {   // vecchia.rs:988-996
        if (range > range_max) {
            const double t = (range - range_max) / range_max;
            return t * t;
        }
        return 0.0;
    }

    // Objective -ll + 1e6 reg(range) (1e12 where the likelihood fails) and,
    // with grads, its gradient in x (vecchia.rs:998-1044); a failed
    // evaluation has gradient 0.
    void run(const std::vector<Vec> &xs, std::vector<double> &f, std::vector<Vec> *grads) {
        const size_t n = xs.size();
        f.assign(n, 1e12);
        if (grads) grads->assign(n, Vec{0.0, 0.0, 0.0});
        std::vector<gs_model> ms;
        std::vector<size_t> at;
        for (size_t i = 0; i < n; i++) {
            gs_model m = model_of(xs[i], kind, param);
            if (!model_ok(m)) continue;
            ms.push_back(m);
            at.push_back(i);
        }
        if (ms.empty()) return;
        const int32_t nm = (int32_t)ms.size();
        std::vector<double> ll((size_t)nm), g((size_t)nm * 3);
        std::vector<int32_t> st((size_t)nm);
        int32_t rc = grads ? gs_vecchia_loglik_grad_batch(plan, values, ms.data(), nm, GS_PREC_FP64,
                                                          ll.data(), g.data(), st.data())
                           : gs_vecchia_loglik_batch(plan, values, ms.data(), nm, GS_PREC_FP64,
                                                     ll.data(), st.data());
        if (rc != GS_OK) {   // candidates are pre-validated: this is a device or size error
            err = rc;
            return;
        }
        for (int32_t j = 0; j < nm; j++) {
            const size_t i = at[(size_t)j];
            if (st[(size_t)j] != 0 || !std::isfinite(ll[(size_t)j])) continue;
            const double psill = ms[(size_t)j].structures[0].sill;
            const double range = ms[(size_t)j].structures[0].range;
            f[i] = -ll[(size_t)j] + 1e6 * reg(range);
            if (grads) {
                Vec &gx = (*grads)[i];
                gx[0] = -g[(size_t)j * 3 + 0] * 2.0 * xs[i][0];
                gx[1] = -g[(size_t)j * 3 + 1] * psill;
                gx[2] = -g[(size_t)j * 3 + 2] * range;
                if (range > range_max)
                    gx[2] += 1e6 * (2.0 * (range - range_max) / (range_max * range_max)) * range;
            }
        }
    }
};

double dot(const Vec &a, const Vec &b) {
    double s = 0.0;
    for (int i = 0; i < kN; i++) s += a[i] * b[i];
    return s;
}

Mat identity() {
    Mat m{};
    for (int i = 0; i < kN; i++) m[i * kN + i] = 1.0;
    return m;
}

// H' = (I - rho s y') H (I - rho y s') + rho s s', rho = 1/(s'y)
// (optim.rs:167-187).
Mat bfgs_update(const Mat &h, const Vec &s, const Vec &y, double sy) {
    const double rho = 1.0 / sy;
    Mat left = identity(), right = identity();
    for (int i = 0; i < kN; i++)
        for (int j = 0; j < kN; j++) {
            left[i * kN + j] -= rho * (s[i] * y[j]);
            right[i * kN + j] -= rho * (y[i] * s[j]);
        }
    auto mm = [](const Mat &a, const Mat &b) {
        Mat c{};
        for (int i = 0; i < kN; i++)
            for (int j = 0; j < kN; j++) {
                double acc = 0.0;
                for (int k = 0; k < kN; k++) acc += a[i * kN + k] * b[k * kN + j];
                c[i * kN + j] = acc;
            }
        return c;
    };
    Mat hn = mm(mm(left, h), right);
    for (int i = 0; i < kN; i++)
        for (int j = 0; j < kN; j++) hn[i * kN + j] += rho * (s[i] * s[j]);
    return hn;
}

// One BFGS run (optim.rs:206-273) as a state machine that asks for one
// (objective, gradient) evaluation at a time.
struct Bfgs {
    enum Phase { kInit, kSearch, kDone } phase = kInit;
    Vec x{}, g{}, p{}, best_x{};
    Mat h = identity();
    double fx = 0.0, best_fx = 0.0, slope = 0.0, alpha = 1.0;
    int iter = 0, max_iter = 200;

    explicit Bfgs(const Vec &x0) : x(x0) {}
    Vec cand() const {
        Vec c = x;
        if (phase == kSearch)
            for (int i = 0; i < kN; i++) c[i] = x[i] + alpha * p[i];
        return c;
    }
    // start of an iteration: stop tests and the search direction
    void begin() {
        for (;;) {
            if (iter >= max_iter || std::sqrt(dot(g, g)) < 1e-10) { phase = kDone; return; }
            for (int i = 0; i < kN; i++) {
                double acc = 0.0;
                for (int j = 0; j < kN; j++) acc += h[i * kN + j] * g[j];
                p[i] = -acc;
            }
            slope = dot(g, p);
            if (slope >= 0.0) {
                h = identity();
                for (int i = 0; i < kN; i++) p[i] = -g[i];
                slope = dot(g, p);
            }
            alpha = 1.0;
            phase = kSearch;
            return;
        }
    }
    void take(const Vec &xe, double fe, const Vec &ge) {
        if (phase == kInit) {
            g = ge;
            fx = fe;
            best_x = x;
            best_fx = fx;
            begin();
            return;
        }
        if (!(fe <= fx + 1e-4 * alpha * slope || alpha < 1e-12)) {
            alpha *= 0.5;
            return;
        }
        Vec s, y;
        for (int i = 0; i < kN; i++) {
            s[i] = xe[i] - x[i];
            y[i] = ge[i] - g[i];
        }
        const double sy = dot(s, y);
        if (sy > 1e-10) h = bfgs_update(h, s, y, sy);
        const bool stalled = std::fabs(fx - fe) <= 1e-13 * (1.0 + std::fabs(fx)) &&
                             std::sqrt(dot(s, s)) <= 1e-10 * (1.0 + std::sqrt(dot(x, x)));
        x = xe;
        g = ge;
        fx = fe;
        if (fx < best_fx) {
            best_fx = fx;
            best_x = x;
        }
        iter++;
        if (stalled) { phase = kDone; return; }
        begin();
    }
};

// One Nelder-Mead run (optim.rs:6-104) as a state machine: each step asks
// for a batch of points (the initial simplex, a reflection, an expansion or
// contraction, or a shrink of n points).
struct NelderMead {
    enum Phase { kInit, kReflect, kExpand, kContract, kShrink, kDone } phase = kInit;
    std::array<Vec, kN + 1> sx{};
    std::array<double, kN + 1> fv{};
    Vec centroid{}, xr{};
    double fr = 0.0;
    int iter = 0, max_iter = 2000;
    double step;

    NelderMead(const Vec &x0, double step_) : step(step_) {
        sx[0] = x0;
        for (int i = 0; i < kN; i++) {
            sx[(size_t)i + 1] = x0;
            sx[(size_t)i + 1][i] += step;
        }
    }
    std::vector<Vec> wants() const {
        switch (phase) {
        case kInit: return std::vector<Vec>(sx.begin(), sx.end());
        case kReflect: return {xr};
        case kExpand: {
            Vec xe;
            for (int j = 0; j < kN; j++) xe[j] = centroid[j] + 2.0 * (centroid[j] - sx[kN][j]);
            return {xe};
        }
        case kContract: {
            Vec xc;
            for (int j = 0; j < kN; j++)
                xc[j] = fr < fv[kN] ? centroid[j] + 0.5 * (xr[j] - centroid[j])
                                    : centroid[j] - 0.5 * (centroid[j] - sx[kN][j]);
            return {xc};
        }
        case kShrink: {
            std::vector<Vec> out;
            for (int i = 1; i <= kN; i++) {
                Vec q;
                for (int j = 0; j < kN; j++) q[j] = sx[0][j] + 0.5 * (sx[(size_t)i][j] - sx[0][j]);
                out.push_back(q);
            }
            return out;
        }
        default: return {};
        }
    }
    // ordering and reflection point of a new iteration
    void begin() {
        if (iter >= max_iter) { phase = kDone; return; }
        std::array<int, kN + 1> idx{};
        for (int i = 0; i <= kN; i++) idx[(size_t)i] = i;
        // stable sort by total order of the objective (optim.rs:26-28)
        for (int i = 1; i <= kN; i++)
            for (int j = i; j > 0; j--) {
                const double a = fv[(size_t)idx[(size_t)j - 1]], b = fv[(size_t)idx[(size_t)j]];
                if (total_key(b) < total_key(a)) std::swap(idx[(size_t)j - 1], idx[(size_t)j]);
                else break;
            }
        std::array<Vec, kN + 1> s2;
        std::array<double, kN + 1> f2;
        for (int i = 0; i <= kN; i++) { s2[(size_t)i] = sx[(size_t)idx[(size_t)i]]; f2[(size_t)i] = fv[(size_t)idx[(size_t)i]]; }
        sx = s2;
        fv = f2;
        if (std::fabs(fv[kN] - fv[0]) <= 1e-12 * std::fmax(std::fabs(fv[0]), 1e-12)) { phase = kDone; return; }
        centroid = Vec{0.0, 0.0, 0.0};
        for (int i = 0; i < kN; i++)
            for (int j = 0; j < kN; j++) centroid[j] += sx[(size_t)i][j] / (double)kN;
        for (int j = 0; j < kN; j++) xr[j] = centroid[j] + 1.0 * (centroid[j] - sx[kN][j]);
        phase = kReflect;
    }
    void take(const std::vector<Vec> &xs, const std::vector<double> &f) {
        switch (phase) {
        case kInit:
            for (int i = 0; i <= kN; i++) fv[(size_t)i] = f[(size_t)i];
            begin();
            return;
        case kReflect:
            fr = f[0];
            if (fr < fv[0]) { phase = kExpand; return; }
            if (fr < fv[kN - 1]) { sx[kN] = xr; fv[kN] = fr; break; }
            phase = kContract;
            return;
        case kExpand:
            if (f[0] < fr) { sx[kN] = xs[0]; fv[kN] = f[0]; }
            else { sx[kN] = xr; fv[kN] = fr; }
            break;
        case kContract:
            if (f[0] < std::fmin(fr, fv[kN])) { sx[kN] = xs[0]; fv[kN] = f[0]; break; }
            phase = kShrink;
            return;
        case kShrink:
            for (int i = 1; i <= kN; i++) { sx[(size_t)i] = xs[(size_t)i - 1]; fv[(size_t)i] = f[(size_t)i - 1]; }
            break;
        default: return;
        }
        iter++;
        begin();
    }
    void result(Vec &x, double &f) const {
        int best = 0;
        for (int i = 1; i <= kN; i++)
            if (fv[(size_t)i] < fv[(size_t)best]) best = i;
        x = sx[(size_t)best];
        f = fv[(size_t)best];
    }
};

// g_and_dg kinds (vecchia.rs:353-385)
bool has_range_derivative(int32_t kind) {
    return kind == GS_SPHERICAL || kind == GS_EXPONENTIAL || kind == GS_GAUSSIAN ||
           kind == GS_MATERN15 || kind == GS_MATERN25;
}

}  // namespace

extern "C" int32_t gs_vecchia_mle(gs_vplan *plan, const double *values, int32_t kind, double param,
                                  gs_model *fit, double *loglik) {
    if (!plan || !values || !fit || !loglik)
        return gs_fail(GS_ERR_INVALID_PARAMETER, "NULL argument");
    // reject_power_kind / reject_invalid_dim_kind (vecchia.rs:946-947)
    if (kind == GS_POWER) return gs_fail(GS_ERR_INVALID_PARAMETER, "Power models cannot be fitted by Vecchia MLE");
    if (kind == GS_CIRCULAR && gs_plan_dim(plan) > 2)
        return gs_fail(GS_ERR_INVALID_PARAMETER, "Circular model is only valid in 1-D and 2-D");
    const int64_t n = gs_plan_points(plan);
    // initial scales (vecchia.rs:958-975)
    double mean = 0.0;
    for (int64_t i = 0; i < n; i++) mean += values[i];
    mean = mean / (double)n;
    double ss = 0.0;
    for (int64_t i = 0; i < n; i++) ss += (values[i] - mean) * (values[i] - mean);
    const double var0 = std::fmax(ss / (double)n, 1e-12);
    const double extent = gs_plan_extent(plan);
    const double range0 = std::fmax(extent / 3.0, 1e-9);
    Evaluator ev{plan, values, kind, param, 10.0 * extent};
    const double lr0 = std::log(range0);
    std::vector<Vec> starts;
    for (double fac : {0.3, 1.0, 3.0})
        starts.push_back(Vec{std::sqrt(0.1 * var0), std::log(0.9 * var0), lr0 + std::log(fac)});

    Vec xb{};
    double fb = 0.0;
    bool have = false;
    auto keep = [&](const Vec &x, double f) {   // first start wins ties (optim.rs:275-288)
        if (!have || f < fb) { xb = x; fb = f; have = true; }
    };
    if (has_range_derivative(kind)) {
        std::vector<Bfgs> runs;
        for (const Vec &x0 : starts) runs.emplace_back(x0);
        for (;;) {
            std::vector<Vec> xs;
            std::vector<size_t> who;
            for (size_t r = 0; r < runs.size(); r++)
                if (runs[r].phase != Bfgs::kDone) { xs.push_back(runs[r].cand()); who.push_back(r); }
            if (xs.empty()) break;
            std::vector<double> f;
            std::vector<Vec> g;
            ev.run(xs, f, &g);
            if (ev.err != GS_OK) return ev.err;
            for (size_t j = 0; j < xs.size(); j++) runs[who[j]].take(xs[j], f[j], g[j]);
        }
        for (auto &r : runs) keep(r.best_x, r.best_fx);
    } else {
        std::vector<NelderMead> runs;
        for (const Vec &x0 : starts) runs.emplace_back(x0, 0.3);
        for (;;) {
            std::vector<Vec> xs;
            std::vector<std::pair<size_t, size_t>> span;   // run, count
            for (size_t r = 0; r < runs.size(); r++) {
                if (runs[r].phase == NelderMead::kDone) continue;
                std::vector<Vec> w = runs[r].wants();
                span.push_back({r, w.size()});
                xs.insert(xs.end(), w.begin(), w.end());
            }
            if (xs.empty()) break;
            std::vector<double> f;
            ev.run(xs, f, nullptr);
            if (ev.err != GS_OK) return ev.err;
            size_t off = 0;
            for (auto &sp : span) {
                std::vector<Vec> xr(xs.begin() + (long)off, xs.begin() + (long)(off + sp.second));
                std::vector<double> fr(f.begin() + (long)off, f.begin() + (long)(off + sp.second));
                runs[sp.first].take(xr, fr);
                off += sp.second;
            }
        }
        for (auto &r : runs) {
            Vec x;
            double f;
            r.result(x, f);
            keep(x, f);
        }
    }
    if (fb >= 1e12)   // vecchia.rs:1055-1063
        return gs_fail(GS_ERR_SINGULAR_SYSTEM,
                                 "Vecchia MLE did not converge to a feasible fit: every starting "
                                 "point stayed on a singular/degenerate covariance");
    *fit = model_of(xb, kind, param);
    *loglik = -fb;
    return GS_OK;
}
