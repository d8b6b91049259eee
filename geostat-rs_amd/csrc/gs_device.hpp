// Device-side covariance model for the gfx950 kernels.
//
// The host resolves a gs_model into a DevModel once per call: per-structure
// inverse range, the anisotropy rotation+scaling matrix (computed on the
// host with the same libm sin/cos as variogram/model.rs:353-367, 407-433,
// instead of re-evaluating trig on every covariance as the reference does),
// and the general-nu Matern normalisation.  Evaluation follows
// VariogramModel::gamma_dh (model.rs:636-646): gamma(0) = 0 exactly, else
// nugget + sum(sill * g(|T dh|, range)); entries are c0 - gamma.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gs {

constexpr int kMaxStructures = 8;

// The chip hands workgroups to its eight XCDs round-robin by workgroup id,
// and each XCD has its own L2.  xcd_block renumbers workgroup b of nb so
// that every XCD takes a contiguous run of the new ids (a bijection on
// [0, nb)): work items in neighbouring positions of a Hilbert-ordered list
// (neighbouring targets, which share most of their neighbours) are then
// gathered through the same L2 instead of eight.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nb) {
    const int64_t x = b & 7, per = nb >> 3, extra = nb & 7;
    return x * per + (x < extra ? x : extra) + (b >> 3);
}

enum Kind : int32_t {
    K_SPH = 0, K_EXP = 1, K_GAU = 2, K_MAT15 = 3, K_MAT25 = 4, K_MAT = 5,
    K_CIR = 6, K_STA = 7, K_HOL = 8, K_WAV = 9, K_POW = 10
};

struct DevStructure {
    int32_t kind;
    int32_t aniso;        // 0 = isotropic, 1 = 2x2 / 3x3 matrix in m
    double param;
    double sill;
    double inv_range;
    double m[9];          // row-major transform: 2-D uses m[0..3]
    double mat_norm;      // 2^(1-nu) / Gamma(nu)   (general Matern)
    double mat_s;         // sqrt(2 nu)
};

// Gauss-Legendre nodes for the general-nu Matern K_nu quadrature
// (variogram/bessel.rs:60-95): 60 nodes on [0,1] and 40 on [1,T].
struct BesselNodes {
    double x60[60], w60[60], x40[40], w40[40];
};

struct DevModel {
    double nugget;
    double c0;            // total sill, VariogramModel::total_sill (model.rs:658-660)
    int32_t ns;
    int32_t pad;
    const BesselNodes *nodes;   // device copy, used by general-nu Matern only
    DevStructure s[kMaxStructures];
};

#if defined(__HIPCC__)
__device__ __forceinline__ double dev_bessel_k(const BesselNodes &gl, double nu, double x) {
    // K_nu(x) = int_0^inf exp(-x cosh t) cosh(nu t) dt (bessel.rs:70-95).
    double integral = 0.0;
    for (int i = 0; i < 60; i++) {
        double t = 0.5 * (gl.x60[i] + 1.0);
        integral += gl.w60[i] * (exp(-x * cosh(t)) * cosh(nu * t));
    }
    integral *= 0.5;
    double xm = x > 1e-12 ? x : 1e-12;
    double a = 60.0 / xm;
    double tm = a < 1.0 ? __builtin_nan("") : log(a + sqrt(a - 1.0) * sqrt(a + 1.0));
    double t_max = fmax(tm, 4.0) + 1.0;
    double integral2 = 0.0;
    for (int i = 0; i < 40; i++) {
        double t = 1.0 + 0.5 * (t_max - 1.0) * (gl.x40[i] + 1.0);
        integral2 += gl.w40[i] * (exp(-x * cosh(t)) * cosh(nu * t));
    }
    integral2 *= 0.5 * (t_max - 1.0);
    return integral + integral2;
}

// Unit-sill shape g(h, range) (model.rs:122-181) in double precision.
__device__ __forceinline__ double shape_g(const DevStructure &s, const BesselNodes *gl, double h) {
    double d = h * s.inv_range;
    switch (s.kind) {
    case K_SPH: return d >= 1.0 ? 1.0 : 1.5 * d - 0.5 * d * d * d;
    case K_EXP: return 1.0 - exp(-d);
    case K_GAU: return 1.0 - exp(-(d * d));
    case K_MAT15: { double q = 1.7320508075688772 * d; return 1.0 - (1.0 + q) * exp(-q); }
    case K_MAT25: { double q = 2.23606797749979 * d; return 1.0 - (1.0 + q + q * q / 3.0) * exp(-q); }
    case K_MAT: {
        if (d <= 0.0) return 0.0;
        double q = s.mat_s * d;
        double corr = s.mat_norm * pow(q, s.param) * dev_bessel_k(*gl, s.param, q);
        corr = corr < 0.0 ? 0.0 : (corr > 1.0 ? 1.0 : corr);
        return 1.0 - corr;
    }
    case K_CIR: {
        double dd = fmin(d, 1.0);
        return 1.0 - 0.63661977236758134 * (acos(dd) - dd * sqrt(1.0 - dd * dd));
    }
    case K_STA: return 1.0 - exp(-pow(d, s.param));
    case K_HOL: return d <= 0.0 ? 0.0 : 1.0 - sin(d) / d;
    case K_WAV: {
        if (d <= 0.0) return 0.0;
        double pd = 3.141592653589793 * d;
        return 1.0 - sin(pd) / pd;
    }
    default: return __builtin_nan("");
    }
}

// Effective isotropic lag |T dh| (model.rs:521-539).
template <int D>
__device__ __forceinline__ double eff_h(const DevStructure &s, const double *dh) {
    double acc = 0.0;
    if (!s.aniso) {
#pragma unroll
        for (int d = 0; d < D; d++) acc += dh[d] * dh[d];
    } else if (D == 2) {
        double e0 = s.m[0] * dh[0] + s.m[1] * dh[1];
        double e1 = s.m[2] * dh[0] + s.m[3] * dh[1];
        acc = e0 * e0 + e1 * e1;
    } else {
        double e0 = s.m[0] * dh[0] + s.m[1] * dh[1] + s.m[2] * dh[2];
        double e1 = s.m[3] * dh[0] + s.m[4] * dh[1] + s.m[5] * dh[2];
        double e2 = s.m[6] * dh[0] + s.m[7] * dh[1] + s.m[8] * dh[2];
        acc = e0 * e0 + e1 * e1 + e2 * e2;
    }
    return sqrt(acc);
}

// sill * g(|T dh| / range) of one structure, kept out of line: the
// libm-grade functions and the Bessel quadrature of shape_g would otherwise
// be inlined at every entry of the unrolled solve kernels, which then need
// far more than 256 VGPRs and spill hundreds of SGPRs.
template <int D>
__device__ __noinline__ double general_term(const DevStructure *s, const BesselNodes *gl,
                                            double dx, double dy, double dz) {
    const double dh[3] = {dx, dy, dz};
    return s->sill * shape_g(*s, gl, eff_h<D>(*s, dh));
}

// ---- lean model class: spherical, exponential, Gaussian, Matern 3/2, 5/2 --
// These five kinds need only exp and sqrt.  Kernels specialised on this
// class use the two routines below instead of the general libm-grade ones,
// which keeps the covariance build short and its register footprint small.

// exp(-s) for s >= 0: Cody-Waite reduction s = k ln2 + r, |r| <= ln2/2,
// then the degree-12 Taylor polynomial of exp(-r) (truncation below 2e-16
// relative on that interval) and a power-of-two scale.  s is clamped at 800
// (exp(-800) underflows to 0 in double).
__device__ __forceinline__ double exp_neg(double s) {
    s = fmin(s, 800.0);
    const double kf = rint(s * 1.4426950408889634);
    double r = fma(-kf, 6.93147180369123816490e-01, s);   // ln2 high part
    r = fma(-kf, 1.90821492927058770002e-10, r);           // ln2 low part
    const double x = -r;
    double p = 2.08767569878680989792e-09;                 // 1/12!
    p = fma(p, x, 2.50521083854417187751e-08);             // 1/11!
    p = fma(p, x, 2.75573192239858906526e-07);             // 1/10!
    p = fma(p, x, 2.75573192239858906526e-06);             // 1/9!
    p = fma(p, x, 2.48015873015873015873e-05);             // 1/8!
    p = fma(p, x, 1.98412698412698412698e-04);             // 1/7!
    p = fma(p, x, 1.38888888888888888889e-03);             // 1/6!
    p = fma(p, x, 8.33333333333333333333e-03);             // 1/5!
    p = fma(p, x, 4.16666666666666666667e-02);             // 1/4!
    p = fma(p, x, 1.66666666666666666667e-01);             // 1/3!
    p = fma(p, x, 0.5);
    p = fma(p, x, 1.0);
    p = fma(p, x, 1.0);
    return ldexp(p, -(int)kf);
}

// sqrt(q) for q >= 0: hardware reciprocal-square-root estimate, then one
// Markstein correction h + (q - h^2) y/2 (about 1 ulp).
__device__ __forceinline__ double sqrt_fast(double q) {
    const double y = __builtin_amdgcn_rsq(q);
    const double h = q * y;
    const double h1 = fma(fma(-h, h, q), 0.5 * y, h);
    return q > 0.0 ? h1 : 0.0;
}

// |T dh|^2 for one structure (model.rs:521-539 before the square root).
template <int D>
__device__ __forceinline__ double eff_h2(const DevStructure &s, const double *dh) {
    if (!s.aniso) {
        double acc = 0.0;
#pragma unroll
        for (int d = 0; d < D; d++) acc += dh[d] * dh[d];
        return acc;
    }
    if (D == 2) {
        double e0 = s.m[0] * dh[0] + s.m[1] * dh[1];
        double e1 = s.m[2] * dh[0] + s.m[3] * dh[1];
        return e0 * e0 + e1 * e1;
    }
    double e0 = s.m[0] * dh[0] + s.m[1] * dh[1] + s.m[2] * dh[2];
    double e1 = s.m[3] * dh[0] + s.m[4] * dh[1] + s.m[5] * dh[2];
    double e2 = s.m[6] * dh[0] + s.m[7] * dh[1] + s.m[8] * dh[2];
    return e0 * e0 + e1 * e1 + e2 * e2;
}

__device__ __forceinline__ double shape_lean(const DevStructure &s, double h2) {
    if (s.kind == K_GAU) {                       // 1 - exp(-(h/a)^2): no root needed
        return 1.0 - exp_neg(h2 * (s.inv_range * s.inv_range));
    }
    const double d = sqrt_fast(h2) * s.inv_range;
    switch (s.kind) {
    case K_SPH: return d >= 1.0 ? 1.0 : 1.5 * d - 0.5 * d * d * d;
    case K_EXP: return 1.0 - exp_neg(d);
    case K_MAT15: { double q = 1.7320508075688772 * d; return 1.0 - (1.0 + q) * exp_neg(q); }
    default: { double q = 2.23606797749979 * d; return 1.0 - (1.0 + q + q * q / 3.0) * exp_neg(q); }
    }
}

// Covariance c0 - gamma(dh) for a separation (the diagonal, where gamma(0) =
// 0 gives exactly c0, is handled by the callers; an all-zero separation to
// a target returns c0 too).  LEAN selects the five-kind fast path.
template <int D, bool LEAN>
__device__ __forceinline__ double cov_offdiag(const DevModel &M, const double *dh) {
    bool allzero = true;
#pragma unroll
    for (int d = 0; d < D; d++) allzero = allzero && (dh[d] == 0.0);
    if (allzero) return M.c0;
    double gam = 0.0;
    for (int i = 0; i < M.ns; i++) {
        if constexpr (LEAN) gam += M.s[i].sill * shape_lean(M.s[i], eff_h2<D>(M.s[i], dh));
        else gam += general_term<D>(&M.s[i], M.nodes, dh[0], D > 1 ? dh[1] : 0.0, D > 2 ? dh[2] : 0.0);
    }
    return M.c0 - (M.nugget + gam);
}

#endif

// True when every structure of the model is in the lean class.
inline bool model_is_lean(const DevModel &M) {
    for (int i = 0; i < M.ns; i++) {
        const int k = M.s[i].kind;
        if (!(k == K_SPH || k == K_EXP || k == K_GAU || k == K_MAT15 || k == K_MAT25)) return false;
    }
    return true;
}

}  // namespace gs
