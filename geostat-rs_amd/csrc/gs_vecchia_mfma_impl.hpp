// loglik_with_plan (vecchia.rs:285-336) for batches of single-structure lean
// models, with the factorization on the matrix cores (FP64 MFMA).
//
// One wavefront handles four ordered points at once, one per MFMA block
// of v_mfma_f64_4x4x4_4b_f64: block b is the 16 lanes l with (l >> 2) & 3
// == b, and lane (x = l >> 4, y = l & 3) of a block holds element [y][x] of
// every 4 x 4 tile of its system.  With that layout the instruction
// computes, per block, out = M2 * M1^T + M3 for the tiles M1 (operand a),
// M2 (operand b) and M3 (operand c), all in the same layout.
//
// A point's system is the bordered matrix of 9 x 9 tiles
//     [ K  . ]     K: conditioning-set covariances (32 x 32, padded rows are
//     [ R  0 ]        c0 on the diagonal and 0 elsewhere)
// where the last tile row R holds the right-hand sides as rows: row 0 the
// covariances to the point (k), row 1 the centred values of the
// conditioning set (z), rows 2-3 zero.  A blocked right-looking Cholesky
// over the eight tile columns of K, carried through R, leaves L_R =
// R L^-T = [ (L^-1 k)^T ; (L^-1 z)^T ] in the last tile row and the
// bottom-right tile at -L_R L_R^T, so mu = (L^-1 k).(L^-1 z) = -tile[0][1]
// and k^T K^-1 k = -tile[0][0] (the reference's w = K^-1 k, mu = w.z,
// var = c0 - w.k).  Per tile column p: the diagonal tile is factored in
// registers (each lane of the block redundantly, from the 16 values read
// back through LDS; its pivots are tested as the reference's Cholesky
// tests them, linalg.rs:135-140), the tiles below it are multiplied by
// L_pp^-T (one MFMA each), and the trailing tiles are updated (one MFMA
// each): 36 + 120 MFMA per point instead of the vector units' column
// steps.  Sums are taken in a fixed order, as in vecchia_kernel.
#pragma once

#include "gs_solve_impl.hpp"

namespace gs {

namespace vm {

constexpr int kTiles = 45;   // 36 of K's lower triangle + 9 of the last tile row

// Per-block LDS strides (doubles), padded so that the four blocks' copies
// start on different banks: the lanes of a wavefront read the same element
// of four blocks at once (coordinates of a tile row or column, the diagonal
// tile, its entries), and unpadded power-of-two strides put all four on one
// bank.
template <int D>
constexpr int kXYS = 32 * D + 8;   // 32 slots of D coordinates, +64 B
constexpr int kTS = 18;            // a 4 x 4 tile, +16 B (16-byte aligned)
constexpr int kDGS = 50;           // 48 diagonal-tile entries, +16 B

__host__ __device__ constexpr int tix(int i, int j) { return i < 8 ? i * (i + 1) / 2 + j : 36 + j; }

// The value, pinned where it is computed: a later select between it and
// another value then stays a select (evaluated in every lane) instead of a
// branch around the evaluation, which would split the code into blocks
// that all carry the tiles.
__device__ __forceinline__ double pin(double v) {
    asm volatile("" : "+v"(v));
    return v;
}

__device__ __forceinline__ double mfma(double a, double b, double c) {
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}

// out = -(M2 * M1^T) + M3: the instruction's negation of operand a (on
// the FP64 matrix instructions the blgp field holds the neg modifiers)
__device__ __forceinline__ double mfma_nega(double a, double b, double c) {
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 1);
}

// 2^(j/256), j = 0..255, rounded to nearest (the exponential's table).
static __constant__ double kExp2Frac256[256] = {
    0x1.0000000000000p+0, 0x1.00b1afa5abcbfp+0, 0x1.0163da9fb3335p+0, 0x1.02168143b0281p+0,
    0x1.02c9a3e778061p+0, 0x1.037d42e11bbccp+0, 0x1.04315e86e7f85p+0, 0x1.04e5f72f654b1p+0,
    0x1.059b0d3158574p+0, 0x1.0650a0e3c1f89p+0, 0x1.0706b29ddf6dep+0, 0x1.07bd42b72a836p+0,
    0x1.0874518759bc8p+0, 0x1.092bdf66607e0p+0, 0x1.09e3ecac6f383p+0, 0x1.0a9c79b1f3919p+0,
    0x1.0b5586cf9890fp+0, 0x1.0c0f145e46c85p+0, 0x1.0cc922b7247f7p+0, 0x1.0d83b23395decp+0,
    0x1.0e3ec32d3d1a2p+0, 0x1.0efa55fdfa9c5p+0, 0x1.0fb66affed31bp+0, 0x1.1073028d7233ep+0,
    0x1.11301d0125b51p+0, 0x1.11edbab5e2ab6p+0, 0x1.12abdc06c31ccp+0, 0x1.136a814f204abp+0,
    0x1.1429aaea92de0p+0, 0x1.14e95934f312ep+0, 0x1.15a98c8a58e51p+0, 0x1.166a45471c3c2p+0,
    0x1.172b83c7d517bp+0, 0x1.17ed48695bbc0p+0, 0x1.18af9388c8deap+0, 0x1.1972658375d2fp+0,
    0x1.1a35beb6fcb75p+0, 0x1.1af99f8138a1cp+0, 0x1.1bbe084045cd4p+0, 0x1.1c82f95281c6bp+0,
    0x1.1d4873168b9aap+0, 0x1.1e0e75eb44027p+0, 0x1.1ed5022fcd91dp+0, 0x1.1f9c18438ce4dp+0,
    0x1.2063b88628cd6p+0, 0x1.212be3578a819p+0, 0x1.21f49917ddc96p+0, 0x1.22bdda27912d1p+0,
    0x1.2387a6e756238p+0, 0x1.2451ffb82140ap+0, 0x1.251ce4fb2a63fp+0, 0x1.25e85711ece75p+0,
    0x1.26b4565e27cddp+0, 0x1.2780e341ddf29p+0, 0x1.284dfe1f56381p+0, 0x1.291ba7591bb70p+0,
    0x1.29e9df51fdee1p+0, 0x1.2ab8a66d10f13p+0, 0x1.2b87fd0dad990p+0, 0x1.2c57e39771b2fp+0,
    0x1.2d285a6e4030bp+0, 0x1.2df961f641589p+0, 0x1.2ecafa93e2f56p+0, 0x1.2f9d24abd886bp+0,
    0x1.306fe0a31b715p+0, 0x1.31432edeeb2fdp+0, 0x1.32170fc4cd831p+0, 0x1.32eb83ba8ea32p+0,
    0x1.33c08b26416ffp+0, 0x1.3496266e3fa2dp+0, 0x1.356c55f929ff1p+0, 0x1.36431a2de883bp+0,
    0x1.371a7373aa9cbp+0, 0x1.37f26231e754ap+0, 0x1.38cae6d05d866p+0, 0x1.39a401b7140efp+0,
    0x1.3a7db34e59ff7p+0, 0x1.3b57fbfec6cf4p+0, 0x1.3c32dc313a8e5p+0, 0x1.3d0e544ede173p+0,
    0x1.3dea64c123422p+0, 0x1.3ec70df1c5175p+0, 0x1.3fa4504ac801cp+0, 0x1.40822c367a024p+0,
    0x1.4160a21f72e2ap+0, 0x1.423fb2709468ap+0, 0x1.431f5d950a897p+0, 0x1.43ffa3f84b9d4p+0,
    0x1.44e086061892dp+0, 0x1.45c2042a7d232p+0, 0x1.46a41ed1d0057p+0, 0x1.4786d668b3237p+0,
    0x1.486a2b5c13cd0p+0, 0x1.494e1e192aed2p+0, 0x1.4a32af0d7d3dep+0, 0x1.4b17dea6db7d7p+0,
    0x1.4bfdad5362a27p+0, 0x1.4ce41b817c114p+0, 0x1.4dcb299fddd0dp+0, 0x1.4eb2d81d8abffp+0,
    0x1.4f9b2769d2ca7p+0, 0x1.508417f4531eep+0, 0x1.516daa2cf6642p+0, 0x1.5257de83f4eefp+0,
    0x1.5342b569d4f82p+0, 0x1.542e2f4f6ad27p+0, 0x1.551a4ca5d920fp+0, 0x1.56070dde910d2p+0,
    0x1.56f4736b527dap+0, 0x1.57e27dbe2c4cfp+0, 0x1.58d12d497c7fdp+0, 0x1.59c0827ff07ccp+0,
    0x1.5ab07dd485429p+0, 0x1.5ba11fba87a03p+0, 0x1.5c9268a5946b7p+0, 0x1.5d84590998b93p+0,
    0x1.5e76f15ad2148p+0, 0x1.5f6a320dceb71p+0, 0x1.605e1b976dc09p+0, 0x1.6152ae6cdf6f4p+0,
    0x1.6247eb03a5585p+0, 0x1.633dd1d1929fdp+0, 0x1.6434634ccc320p+0, 0x1.652b9febc8fb7p+0,
    0x1.6623882552225p+0, 0x1.671c1c70833f6p+0, 0x1.68155d44ca973p+0, 0x1.690f4b19e9538p+0,
    0x1.6a09e667f3bcdp+0, 0x1.6b052fa75173ep+0, 0x1.6c012750bdabfp+0, 0x1.6cfdcddd47645p+0,
    0x1.6dfb23c651a2fp+0, 0x1.6ef9298593ae5p+0, 0x1.6ff7df9519484p+0, 0x1.70f7466f42e87p+0,
    0x1.71f75e8ec5f74p+0, 0x1.72f8286ead08ap+0, 0x1.73f9a48a58174p+0, 0x1.74fbd35d7cbfdp+0,
    0x1.75feb564267c9p+0, 0x1.77024b1ab6e09p+0, 0x1.780694fde5d3fp+0, 0x1.790b938ac1cf6p+0,
    0x1.7a11473eb0187p+0, 0x1.7b17b0976cfdbp+0, 0x1.7c1ed0130c132p+0, 0x1.7d26a62ff86f0p+0,
    0x1.7e2f336cf4e62p+0, 0x1.7f3878491c491p+0, 0x1.80427543e1a12p+0, 0x1.814d2add106d9p+0,
    0x1.82589994cce13p+0, 0x1.8364c1eb941f7p+0, 0x1.8471a4623c7adp+0, 0x1.857f4179f5b21p+0,
    0x1.868d99b4492edp+0, 0x1.879cad931a436p+0, 0x1.88ac7d98a6699p+0, 0x1.89bd0a478580fp+0,
    0x1.8ace5422aa0dbp+0, 0x1.8be05bad61778p+0, 0x1.8cf3216b5448cp+0, 0x1.8e06a5e0866d9p+0,
    0x1.8f1ae99157736p+0, 0x1.902fed0282c8ap+0, 0x1.9145b0b91ffc6p+0, 0x1.925c353aa2fe2p+0,
    0x1.93737b0cdc5e5p+0, 0x1.948b82b5f98e5p+0, 0x1.95a44cbc8520fp+0, 0x1.96bdd9a7670b3p+0,
    0x1.97d829fde4e50p+0, 0x1.98f33e47a22a2p+0, 0x1.9a0f170ca07bap+0, 0x1.9b2bb4d53fe0dp+0,
    0x1.9c49182a3f090p+0, 0x1.9d674194bb8d5p+0, 0x1.9e86319e32323p+0, 0x1.9fa5e8d07f29ep+0,
    0x1.a0c667b5de565p+0, 0x1.a1e7aed8eb8bbp+0, 0x1.a309bec4a2d33p+0, 0x1.a42c980460ad8p+0,
    0x1.a5503b23e255dp+0, 0x1.a674a8af46052p+0, 0x1.a799e1330b358p+0, 0x1.a8bfe53c12e59p+0,
    0x1.a9e6b5579fdbfp+0, 0x1.ab0e521356ebap+0, 0x1.ac36bbfd3f37ap+0, 0x1.ad5ff3a3c2774p+0,
    0x1.ae89f995ad3adp+0, 0x1.afb4ce622f2ffp+0, 0x1.b0e07298db666p+0, 0x1.b20ce6c9a8952p+0,
    0x1.b33a2b84f15fbp+0, 0x1.b468415b749b1p+0, 0x1.b59728de5593ap+0, 0x1.b6c6e29f1c52ap+0,
    0x1.b7f76f2fb5e47p+0, 0x1.b928cf22749e4p+0, 0x1.ba5b030a1064ap+0, 0x1.bb8e0b79a6f1fp+0,
    0x1.bcc1e904bc1d2p+0, 0x1.bdf69c3f3a207p+0, 0x1.bf2c25bd71e09p+0, 0x1.c06286141b33dp+0,
    0x1.c199bdd85529cp+0, 0x1.c2d1cd9fa652cp+0, 0x1.c40ab5fffd07ap+0, 0x1.c544778fafb22p+0,
    0x1.c67f12e57d14bp+0, 0x1.c7ba88988c933p+0, 0x1.c8f6d9406e7b5p+0, 0x1.ca3405751c4dbp+0,
    0x1.cb720dcef9069p+0, 0x1.ccb0f2e6d1675p+0, 0x1.cdf0b555dc3fap+0, 0x1.cf3155b5bab74p+0,
    0x1.d072d4a07897cp+0, 0x1.d1b532b08c968p+0, 0x1.d2f87080d89f2p+0, 0x1.d43c8eacaa1d6p+0,
    0x1.d5818dcfba487p+0, 0x1.d6c76e862e6d3p+0, 0x1.d80e316c98398p+0, 0x1.d955d71ff6075p+0,
    0x1.da9e603db3285p+0, 0x1.dbe7cd63a8315p+0, 0x1.dd321f301b460p+0, 0x1.de7d5641c0658p+0,
    0x1.dfc97337b9b5fp+0, 0x1.e11676b197d17p+0, 0x1.e264614f5a129p+0, 0x1.e3b333b16ee12p+0,
    0x1.e502ee78b3ff6p+0, 0x1.e653924676d76p+0, 0x1.e7a51fbc74c83p+0, 0x1.e8f7977cdb740p+0,
    0x1.ea4afa2a490dap+0, 0x1.eb9f4867cca6ep+0, 0x1.ecf482d8e67f1p+0, 0x1.ee4aaa2188510p+0,
    0x1.efa1bee615a27p+0, 0x1.f0f9c1cb6412ap+0, 0x1.f252b376bba97p+0, 0x1.f3ac948dd7274p+0,
    0x1.f50765b6e4540p+0, 0x1.f6632798844f8p+0, 0x1.f7bfdad9cbe14p+0, 0x1.f91d802243c89p+0,
    0x1.fa7c1819e90d8p+0, 0x1.fbdba3692d514p+0, 0x1.fd3c22b8f71f1p+0, 0x1.fe9d96b2a23d9p+0,
};
static __shared__ double g_exp256[256];

__device__ __forceinline__ void load_exp256() {
    for (int i = threadIdx.x; i < 256; i += blockDim.x) g_exp256[i] = kExp2Frac256[i];
    __syncthreads();
}

// e^-s for s >= 0 (s clamped at 800, past which the result is 0).  With
// k = round(-s 256 / ln2) = 256 n + j (0 <= j < 256) and r = -s - k ln2/256
// (|r| <= ln2/512, the two-part constant of ln2/256), e^-s = 2^n 2^(j/256)
// e^r, and e^r - 1 = r + r^2/2 + r^3/6 + r^4/24 to within 4e-17.  k comes
// from the 1.5 * 2^52 shifter: round to nearest leaves it in the low word.
__device__ __forceinline__ double exp_neg256(double s) {
    s = fmin(s, 800.0);
    const double kb = fma(s, -0x1.71547652b82fep+8, 0x1.8p52);
    const double kf = kb - 0x1.8p52;
    const int k = (int)(uint32_t)__builtin_bit_cast(uint64_t, kb);
    double r = fma(kf, -0x1.62e42fee00000p-9, -s);
    r = fma(kf, -0x1.a39ef35793c76p-41, r);
    double p = fma(r, 1.0 / 24.0, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    const double q = p * r;
    const double t = g_exp256[k & 255];
    return ldexp(fma(t, q, t), k >> 8);
}

// One lean structure (KIND: 0 spherical, 1 Matern 3/2, 2 Matern 5/2, 3
// exponential, 4 Gaussian) in its scaled frame: a point c (relative to a
// centre) maps to q = m c, where m is the structure's rotation / scaling
// over the range, times sqrt(3) (Matern 3/2) or sqrt(5) (Matern 5/2), so
// that the kind's argument is |q_a - q_b| (|.|^2 for the Gaussian) and the
// sill is folded into the coefficients:
//   spherical    d < 1: sill - d (1.5 sill - 0.5 sill d^2), else 0
//   Matern 3/2   (sill + sill x) e^-x
//   Matern 5/2   (sill + sill x + sill/3 x^2) e^-x
//   exponential  sill e^-x;  Gaussian  sill e^-(x^2)
// (model.rs:122-181: C = sill (1 - g); the same values as lean_term up to
// rounding.)
template <int D, bool ANISO, int KIND>
struct Frame {
    double m[ANISO ? D * D : 1];
    double a0, a1, a2;
};

template <int D, bool ANISO, int KIND>
__device__ __forceinline__ Frame<D, ANISO, KIND> make_frame(const DevStructure &S) {
    Frame<D, ANISO, KIND> F;
    const double f = KIND == 1 ? 1.7320508075688772 : (KIND == 2 ? 2.23606797749979 : 1.0);
    const double sc = S.inv_range * f;
    if constexpr (ANISO) {
#pragma unroll
        for (int i = 0; i < D * D; i++) F.m[i] = S.m[i] * sc;
    } else {
        F.m[0] = sc;
    }
    const double sill = S.sill;
    F.a0 = sill;
    F.a1 = KIND == 0 ? 1.5 * sill : sill;
    F.a2 = KIND == 0 ? 0.5 * sill : sill * (1.0 / 3.0);
    return F;
}

template <int D, bool ANISO, int KIND>
__device__ __forceinline__ void to_frame(const Frame<D, ANISO, KIND> &F, const double *c, double *q) {
#pragma unroll
    for (int k = 0; k < D; k++) {
        if constexpr (ANISO) {
            double e = 0.0;
#pragma unroll
            for (int d = 0; d < D; d++) e = fma(F.m[k * D + d], c[d], e);
            q[k] = e;
        } else {
            q[k] = F.m[0] * c[k];
        }
    }
}

// The covariance from the squared frame distance h2.
template <int D, bool ANISO, int KIND>
__device__ __forceinline__ double cov_h2(const Frame<D, ANISO, KIND> &F, double h2) {
    if constexpr (KIND == 4) return F.a0 * exp_neg256(h2);
    const double d = sqrt_lean(h2);
    if constexpr (KIND == 0) return d >= 1.0 ? 0.0 : fma(-d, fma(-F.a2, h2, F.a1), F.a0);
    if constexpr (KIND == 1) return fma(F.a1, d, F.a0) * exp_neg256(d);
    if constexpr (KIND == 2) return fma(fma(F.a2, d, F.a1), d, F.a0) * exp_neg256(d);
    return F.a0 * exp_neg256(d);
}

// Empty slots of a system sit far out on the first axis of the frame, slot
// j at kFar (j + 1) range units from the centre: at least 10^30 ranges from
// each other and from every datum, so their covariances with everything
// evaluate to exactly zero (the exponential underflows; the spherical kind
// is zero past its range) and no entry needs a mask.  Frame units keep the
// squared distances finite whatever the range.
constexpr double kFar = 1e30;

// Covariance of two distinct points from their frame coordinates.  The
// handles reject duplicate coordinates (kriging.rs:312, vecchia.rs:248),
// so two data points are never at zero separation.
template <int D, bool ANISO, int KIND>
__device__ __forceinline__ double cov_q(const Frame<D, ANISO, KIND> &F, const double *qa, const double *qb) {
    double h2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; d++) {
        const double e = qa[d] - qb[d];
        h2 = fma(e, e, h2);
    }
    return cov_h2(F, h2);
}

// Element [y][x] of L^-1 for the diagonal tile in tt (row-major 4 x 4,
// written by the block): every lane of the block factors the tile
// redundantly and forms column x of the factor's inverse, then keeps entry
// y (selects between values every lane holds, no lane-dependent branch).  ok is
// cleared when a pivot fails the singular test (GE: pivot >= thresh, else
// pivot > thresh).
template <bool GE>
__device__ __forceinline__ double tile_linv(const double *tt, int y, int x, double thresh, bool &ok) {
    const double m00 = tt[0];
    const double m10 = tt[4], m11 = tt[5];
    const double m20 = tt[8], m21 = tt[9], m22 = tt[10];
    const double m30 = tt[12], m31 = tt[13], m32 = tt[14], m33 = tt[15];
    const double d0 = m00;
    const double i0 = rsqrt_nr(d0);
    const double l10 = m10 * i0, l20 = m20 * i0, l30 = m30 * i0;
    const double d1 = m11 - l10 * l10;
    const double i1 = rsqrt_nr(d1);
    const double l21 = (m21 - l20 * l10) * i1, l31 = (m31 - l30 * l10) * i1;
    const double d2 = (m22 - l20 * l20) - l21 * l21;
    const double i2 = rsqrt_nr(d2);
    const double l32 = ((m32 - l30 * l20) - l31 * l21) * i2;
    const double d3 = ((m33 - l30 * l30) - l31 * l31) - l32 * l32;
    const double i3 = rsqrt_nr(d3);
    if constexpr (GE) ok = ok && d0 >= thresh && d1 >= thresh && d2 >= thresh && d3 >= thresh;
    else ok = ok && d0 > thresh && d1 > thresh && d2 > thresh && d3 > thresh;
    // column x of L^-1 by forward substitution, L u = e_x: zero above the
    // diagonal, 1 / L_xx on it, and below it the products of the closed
    // form (for x = 0: u1 = -l10 i0 i1, u2 = -(l20 i0 + l21 u1) i2, ...).
    // Pinned: computed in every lane before the selects, not sunk into
    // lane-dependent branches.
    const double u0 = x == 0 ? pin(i0) : 0.0;
    const double u1 = x == 1 ? pin(i1) : pin(-l10 * u0 * i1);
    const double u2 = x == 2 ? pin(i2) : pin(-(l20 * u0 + l21 * u1) * i2);
    const double u3 = x == 3 ? pin(i3) : pin(-((l30 * u0 + l31 * u1) + l32 * u2) * i3);
    double r = y == 1 ? u1 : u0;
    r = y == 2 ? u2 : r;
    return y == 3 ? u3 : r;
}

// Blocked right-looking Cholesky of the 9 x 9 tile system over its eight
// tile columns (the last tile row carried along).  Per column q: the tiles
// below the diagonal are multiplied by L_qq^-T (one MFMA each) and the
// trailing tiles updated (one MFMA each).  The next diagonal tile is
// updated first and written to tt before the other trailing updates are
// issued, so its LDS round trip runs under them.
template <bool GE>
__device__ __forceinline__ void factor_tiles(double (&X)[kTiles], double *tt, int y, int x,
                                             double thresh, bool &ok) {
    tt[4 * y + x] = X[tix(0, 0)];
    wave_sync();
    double linv = tile_linv<GE>(tt, y, x, thresh, ok);
    static_for<0, 8>([&](auto PP) {
        constexpr int q = decltype(PP)::value;
        // tiles below: L_iq = X_iq L_qq^-T
        static_for<q + 1, 9>([&](auto II) {
            constexpr int i = decltype(II)::value;
            X[tix(i, q)] = mfma(linv, X[tix(i, q)], 0.0);
        });
        if constexpr (q < 7) {
            X[tix(q + 1, q + 1)] = mfma_nega(X[tix(q + 1, q)], X[tix(q + 1, q)], X[tix(q + 1, q + 1)]);
            wave_sync();
            tt[4 * y + x] = X[tix(q + 1, q + 1)];
        }
        // trailing tiles: X_ij -= L_iq L_jq^T
        static_for<q + 1, 9>([&](auto II) {
            constexpr int i = decltype(II)::value;
            static_for<q + 1, i + 1>([&](auto JJ) {
                constexpr int j = decltype(JJ)::value;
                if constexpr (!(i == q + 1 && j == q + 1 && q < 7))
                    X[tix(i, j)] = mfma_nega(X[tix(j, q)], X[tix(i, q)], X[tix(i, j)]);
            });
        });
        if constexpr (q < 7) {
            wave_sync();
            linv = tile_linv<GE>(tt, y, x, thresh, ok);
        }
    });
}

}  // namespace vm

template <int D, bool ANISO, int KIND>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(3))) void vecchia_mfma_kernel(VecchiaParams P) {
    using namespace vm;
    constexpr int NP = 32;                    // conditioning slots per point
    __shared__ double s_xy[4 * kXYS<D>];      // per block: the conditioning set relative to the point
    __shared__ double s_q[4 * kXYS<D>];       // per block: the same in the current model's frame
    __shared__ double s_rhs[4 * 2 * NP + NP]; // per block: the last tile row's rows k, z; a zero row
    __shared__ double s_t[4 * kTS];
    __shared__ double s_dg[4 * kDGS];         // per block: the diagonal tiles' entries below the diagonal
    __shared__ double s_sum[4 * kMaxModelsPerLaunch];
    const int lane = threadIdx.x;
    const int blk = (lane >> 2) & 3;          // the point of this lane
    const int x_ = lane >> 4, y_ = lane & 3;  // element [y][x] of every tile
    double *xy = s_xy + blk * kXYS<D>;
    double *qs = s_q + blk * kXYS<D>;
    double *rhs = s_rhs + blk * 2 * NP;
    double *ks = rhs;                         // row 0: covariances to the point
    double *zs = rhs + NP;                    // row 1: centred values of the conditioning set
    // the lane's row of the last tile row (rows 2 and 3 are zero)
    const double *rrow = y_ < 2 ? rhs + y_ * NP : s_rhs + 8 * NP;
    double *tt = s_t + blk * kTS;
    double *dgs = s_dg + blk * kDGS;
    const int g = x_ * 4 + y_;                // rank of the lane inside its block
    const int nm = P.n_models;
    for (int p = g; p < nm; p += 16) s_sum[blk * kMaxModelsPerLaunch + p] = 0.0;
    if (lane < NP) s_rhs[8 * NP + lane] = 0.0;
    load_exp256();
    wave_sync();
    const double ln2pi = 1.8378770664093453;  // ln(2 pi)

    const int64_t k0 = (int64_t)blockIdx.x * P.points_per_block;
    int64_t k1 = k0 + P.points_per_block;
    if (k1 > P.n) k1 = P.n;
    // every lane runs every step (the matrix instructions take the whole
    // wavefront); a block past the end computes an empty system and drops it
    for (int64_t kb = k0; kb < k1; kb += 4) {
        const int64_t k = kb + blk;
        const bool live_pt = k < k1;
        const int64_t kk = live_pt ? k : k0;
        const int64_t self = P.order[kk];
        const int cnt_ = live_pt ? P.nb_count[kk] : 0;
        const int32_t *nbl = P.nb + kk * (int64_t)P.m;
        double tg[D];
#pragma unroll
        for (int d = 0; d < D; d++) tg[d] = P.coords[self * D + d];
        const double zi = P.zc[self];
        // the conditioning set relative to the point (empty slots at the point)
        for (int j = g; j < NP; j += 16) {
            const int64_t id = j < cnt_ ? nbl[j] : 0;
#pragma unroll
            for (int d = 0; d < D; d++) xy[j * D + d] = j < cnt_ ? P.coords[id * D + d] - tg[d] : 0.0;
            zs[j] = j < cnt_ ? P.zc[id] : 0.0;
        }
        wave_sync();
        double bvar = 1.0, brr = 0.0;
        for (int p = 0; p < nm; p++) {
            const DevModel &Mdl = P.models[p];
            const double c0 = Mdl.c0;
            const Frame<D, ANISO, KIND> F = make_frame<D, ANISO, KIND>(Mdl.s[0]);
            const double thresh = 1e-12 * c0;
            // the set in this model's frame; empty slots far out
            for (int j = g; j < NP; j += 16) {
                double c[D], q[D];
#pragma unroll
                for (int d = 0; d < D; d++) c[d] = xy[j * D + d];
                to_frame(F, c, q);
#pragma unroll
                for (int d = 0; d < D; d++)
                    qs[j * D + d] = j < cnt_ ? q[d] : (d == 0 ? kFar * (j + 1) : 0.0);
            }
            wave_sync();
            // lane coordinates made opaque per model: the coordinates of the
            // build are then read inside the model loop instead of being
            // hoisted out of it into long-lived registers
            const int y = opaque(y_), x = opaque(x_);
            double X[kTiles];
            // ---- build: element [y][x] of tile (i, j) is row 4i+y, column 4j+x.
            // The 48 distinct entries of the diagonal tiles (below their
            // diagonals) and the 32 covariances to the point are spread over
            // the 16 lanes of the block (3 + 2 each) and exchanged through
            // LDS; every lane evaluates one entry of each of the 28 tiles
            // below the diagonal.
#pragma unroll
            for (int e3 = 0; e3 < 3; e3++) {
                const int u = 3 * g + e3;              // diagonal tile u / 6, entry u % 6
                const int t = u / 6, e = u % 6;
                const int r = e >= 3 ? 3 : (e >= 1 ? 2 : 1);
                const int c = e - r * (r - 1) / 2;
                const int a = 4 * t + r, b = 4 * t + c;
                dgs[u] = cov_q(F, qs + a * D, qs + b * D);
            }
            // to the point, at the origin (it is not in its own conditioning set)
#pragma unroll
            for (int e2 = 0; e2 < 2; e2++) {
                const int b = 2 * g + e2;
                double h2 = 0.0;
#pragma unroll
                for (int d = 0; d < D; d++) h2 = fma(qs[b * D + d], qs[b * D + d], h2);
                ks[b] = cov_h2(F, h2);
            }
            static_for<1, 8>([&](auto II) {
                constexpr int i = decltype(II)::value;
                const int a = 4 * i + y;
                double pa[D];
#pragma unroll
                for (int d = 0; d < D; d++) pa[d] = qs[a * D + d];
                static_for<0, i>([&](auto JJ) {
                    constexpr int j = decltype(JJ)::value;
                    const int b = 4 * j + x;
                    double pb[D];
#pragma unroll
                    for (int d = 0; d < D; d++) pb[d] = qs[b * D + d];
                    X[tix(i, j)] = pin(cov_q(F, pa, pb));
                    if constexpr ((tix(i, j) % 2) == 1) __builtin_amdgcn_sched_barrier(0);
                });
            });
            wave_sync();
            {
                const int hi = y > x ? y : x, lo = y > x ? x : y;
                const int e = hi * (hi - 1) / 2 + lo;  // entry of (hi, lo) below the diagonal
                static_for<0, 8>([&](auto TT) {
                    constexpr int t = decltype(TT)::value;
                    X[tix(t, t)] = y == x ? c0 : dgs[6 * t + e];
                });
            }
            static_for<0, 8>([&](auto JJ) {
                constexpr int j = decltype(JJ)::value;
                X[tix(8, j)] = rrow[4 * j + x];
            });
            X[tix(8, 8)] = 0.0;
            wave_sync();
            // ---- blocked Cholesky over the eight tile columns of K
            bool ok = true;
            factor_tiles<false>(X, tt, y, x, thresh, ok);
            // bottom-right tile: [0][0] = -|L^-1 k|^2 (lane x = 0, y = 0),
            // [0][1] = -(L^-1 k).(L^-1 z) (lane x = 1, y = 0)
            const double t00 = __shfl(X[tix(8, 8)], 4 * blk);
            const double t01 = __shfl(X[tix(8, 8)], 16 + 4 * blk);
            // conditional variance and residual (mu = -t01; a point without
            // predecessors has the marginal c0); lane p mod 16 of the block
            // keeps them, and every 16 models each lane turns its pair into
            // its model's term: one log and one division per 16 models
            const double var = cnt_ > 0 ? fmax(c0 + t00, 1e-12) : c0;
            const double rr = cnt_ > 0 ? zi - (-t01) : zi;
            const bool mine = g == (p & 15);
            bvar = mine ? var : bvar;
            brr = mine ? rr : brr;
            if (cnt_ > 0 && g == 0 && live_pt && !ok) atomicMax(&P.status[p], 3);
            if ((p & 15) == 15 || p == nm - 1) {
                const int pm = (p & ~15) + g;
                if (live_pt && pm <= p)
                    s_sum[blk * kMaxModelsPerLaunch + pm] += -0.5 * (ln2pi + log(bvar) + brr * brr / bvar);
            }
        }
        wave_sync();
    }
    __syncthreads();
    for (int p = lane; p < nm; p += kWave) {
        double s = 0.0;
        for (int q = 0; q < 4; q++) s += s_sum[q * kMaxModelsPerLaunch + p];
        P.partial[(int64_t)blockIdx.x * nm + p] = s;
    }
}

// Kriging::predict_inner (kriging.rs:569-644) for simple (M = 0) and
// ordinary (M = 1) kriging with 17-32 neighbours and a single lean
// structure, on the matrix cores: the system of vecchia_mfma_kernel with
// the last tile row holding f (the constant drift column), k (covariances
// to the target) and z (the values; minus the mean for simple kriging).
// The elimination leaves the Gram matrix of a = L^-1 f, b = L^-1 k and
// c = L^-1 z in the bottom-right tile, and the bordered solve of the
// reference follows from it: for ordinary kriging (one drift term, f(t) =
// 1) the Lagrange multiplier is (a.b - 1) / a.a, the value b.c - (a.c) mu
// and the variance c0 - b.b + (a.b - 1)^2 / a.a; for simple kriging the
// value is mean + b.c and the variance c0 - b.b.  Singular systems as in
// krige_fp64_kernel: Cholesky pivots below 1e-12 of the bordered matrix's
// scale, or a drift Schur complement a.a below it (linalg.rs:36-57).
template <int D, bool ANISO, int KIND, int M>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(3))) void krige_mfma_kernel(SolveParams P) {
    using namespace vm;
    static_assert(M == 0 || M == 1, "simple or ordinary kriging");
    constexpr int NP = 32;
    __shared__ double s_xy[4 * kXYS<D>];      // per block: the neighbours in the frame, target at the origin
    __shared__ double s_rhs[4 * 3 * NP + NP]; // per block: the last tile row's rows f, k, z; a zero row
    __shared__ double s_t[4 * kTS];
    __shared__ double s_dg[4 * kDGS];
    const int lane = threadIdx.x;
    const int blk = (lane >> 2) & 3;
    const int x_ = lane >> 4, y_ = lane & 3;
    const int g = x_ * 4 + y_;
    double *xy = s_xy + blk * kXYS<D>;
    double *rhs = s_rhs + blk * 3 * NP;
    double *fs = rhs, *ks = rhs + NP, *vs = rhs + 2 * NP;
    const double *rrow = y_ < 3 ? rhs + y_ * NP : s_rhs + 12 * NP;
    double *tt = s_t + blk * kTS;
    double *dgs = s_dg + blk * kDGS;
    const DevModel &Mdl = *P.model;
    const double c0 = Mdl.c0;
    const Frame<D, ANISO, KIND> F = make_frame<D, ANISO, KIND>(Mdl.s[0]);
    const double thresh = 1e-12 * (M >= 1 ? fmax(c0, 1.0) : c0);
    if (lane < NP) s_rhs[12 * NP + lane] = 0.0;
    load_exp256();
    wave_sync();
    const int64_t nsys = P.list ? (int64_t)(*P.list_len) : P.T;
    for (int64_t sb = xcd_block(blockIdx.x, gridDim.x) * 4; sb < nsys; sb += (int64_t)gridDim.x * 4) {
        const int64_t sys = sb + blk;
        const bool live = sys < nsys;
        const int64_t t = P.list ? (int64_t)P.list[live ? sys : sb] : (live ? sys : sb);
        const int craw = P.count[t];
        const bool bad = craw < P.min_nb || craw < M || craw > NP;
        const int cnt_ = live && !bad ? craw : 0;
        double tg[D];
#pragma unroll
        for (int d = 0; d < D; d++) tg[d] = P.targets[t * D + d];
        const int32_t *nbl = P.nbr + t * (int64_t)P.stride;
        // the neighbours relative to the target, in the frame; a neighbour
        // on the target maps to the origin, any other to a nonzero point
        // (an image that underflows to zero is moved off the origin); empty
        // slots far out
        for (int j = g; j < NP; j += 16) {
            const int64_t id = j < cnt_ ? nbl[j] : 0;
            double c[D], q[D];
            bool on = true, qz = true;
#pragma unroll
            for (int d = 0; d < D; d++) {
                c[d] = j < cnt_ ? P.coords[id * D + d] - tg[d] : 0.0;
                on = on && c[d] == 0.0;
            }
            to_frame(F, c, q);
#pragma unroll
            for (int d = 0; d < D; d++) qz = qz && q[d] == 0.0;
            if (qz && !on) q[0] = 0x1p-1074;
            if (j >= cnt_) {
#pragma unroll
                for (int d = 0; d < D; d++) q[d] = d == 0 ? kFar * (j + 1) : 0.0;
            }
#pragma unroll
            for (int d = 0; d < D; d++) xy[j * D + d] = q[d];
            vs[j] = j < cnt_ ? (M == 0 ? P.values[id] - P.sk_mean : P.values[id]) : 0.0;
            fs[j] = (M == 1 && j < cnt_) ? 1.0 : 0.0;
        }
        wave_sync();
        const int y = opaque(y_), x = opaque(x_);
        double X[kTiles];
#pragma unroll
        for (int e3 = 0; e3 < 3; e3++) {
            const int u = 3 * g + e3;
            const int tt6 = u / 6, e = u % 6;
            const int r = e >= 3 ? 3 : (e >= 1 ? 2 : 1);
            const int c = e - r * (r - 1) / 2;
            const int a = 4 * tt6 + r, b = 4 * tt6 + c;
            dgs[u] = cov_q(F, xy + a * D, xy + b * D);
        }
        // to the target, at the origin: c0 for a neighbour on it (gamma(0) = 0)
#pragma unroll
        for (int e2 = 0; e2 < 2; e2++) {
            const int b = 2 * g + e2;
            double h2 = 0.0;
            bool on = true;
#pragma unroll
            for (int d = 0; d < D; d++) {
                h2 = fma(xy[b * D + d], xy[b * D + d], h2);
                on = on && xy[b * D + d] == 0.0;
            }
            const double v = cov_h2(F, h2);
            ks[b] = on ? c0 : v;
        }
        static_for<1, 8>([&](auto II) {
            constexpr int i = decltype(II)::value;
            const int a = 4 * i + y;
            double pa[D];
#pragma unroll
            for (int d = 0; d < D; d++) pa[d] = xy[a * D + d];
            static_for<0, i>([&](auto JJ) {
                constexpr int j = decltype(JJ)::value;
                const int b = 4 * j + x;
                double pb[D];
#pragma unroll
                for (int d = 0; d < D; d++) pb[d] = xy[b * D + d];
                X[tix(i, j)] = pin(cov_q(F, pa, pb));
                if constexpr ((tix(i, j) % 2) == 1) __builtin_amdgcn_sched_barrier(0);
            });
        });
        wave_sync();
        {
            const int hi = y > x ? y : x, lo = y > x ? x : y;
            const int e = hi * (hi - 1) / 2 + lo;
            static_for<0, 8>([&](auto TT) {
                constexpr int q = decltype(TT)::value;
                X[tix(q, q)] = y == x ? c0 : dgs[6 * q + e];
            });
        }
        static_for<0, 8>([&](auto JJ) {
            constexpr int j = decltype(JJ)::value;
            X[tix(8, j)] = rrow[4 * j + x];
        });
        X[tix(8, 8)] = 0.0;
        wave_sync();
        bool ok = true;
        factor_tiles<true>(X, tt, y, x, thresh, ok);
        // Gram entries (negated) of a, b, c: element [r][s] at lane x = s, y = r
        const double gaa = -__shfl(X[tix(8, 8)], 4 * blk);
        const double gab = -__shfl(X[tix(8, 8)], 16 + 4 * blk);
        const double gac = -__shfl(X[tix(8, 8)], 32 + 4 * blk);
        const double gbb = -__shfl(X[tix(8, 8)], 16 + 4 * blk + 1);
        const double gbc = -__shfl(X[tix(8, 8)], 32 + 4 * blk + 1);
        if (live && g == 0) {
            double value, var, lag;
            if constexpr (M == 0) {
                value = P.sk_mean + gbc;
                var = c0 - gbb;
                lag = __builtin_nan("");
            } else {
                ok = ok && gaa > 0.0 && !(gaa < thresh);
                lag = (gab - 1.0) / gaa;
                value = gbc - gac * lag;
                var = (c0 - gbb) + (gab - 1.0) * lag;
            }
            const bool good = !bad && ok;
            P.value[t] = good ? value : __builtin_nan("");
            P.variance[t] = good ? fmax(var, 0.0) : __builtin_nan("");
            if (P.lagrange) P.lagrange[t] = good ? lag : __builtin_nan("");
            if (P.status) P.status[t] = bad ? (craw > NP ? 4 : (craw < P.min_nb ? 5 : 1)) : (ok ? 0 : 3);
        }
        wave_sync();
    }
}

template <int D, bool ANISO, int KIND, int M>
hipError_t launch_krige_mfma_inst(const SolveParams &p, int64_t grid_groups, hipStream_t s) {
    int64_t blocks = (grid_groups + 3) / 4;
    if (blocks < 1) blocks = 1;
    if (blocks > (1 << 30)) blocks = 1 << 30;
    hipLaunchKernelGGL((krige_mfma_kernel<D, ANISO, KIND, M>), dim3((unsigned)blocks), dim3(kWave), 0,
                       s, p);
    return hipGetLastError();
}

template <int D, bool ANISO, int KIND>
hipError_t launch_vecchia_mfma_inst(const VecchiaParams &p, int64_t n_blocks, hipStream_t s) {
    hipLaunchKernelGGL((vecchia_mfma_kernel<D, ANISO, KIND>), dim3((unsigned)n_blocks), dim3(kWave), 0,
                       s, p);
    return hipGetLastError();
}

}  // namespace gs
