// Batched kriging / Vecchia solves on gfx950 (CDNA4).
//
// One kriging system per group of G lanes inside a 64-wide wavefront.  The
// data-data covariance block C (n <= NB) is distributed by rows over the
// group: lane g owns rows i = G*r + g (r < R = NB/G) and keeps, for each
// owned row, the columns up to the end of its row block, entirely in VGPRs.
// A right-looking Cholesky runs with the pivot broadcast inside the group by
// DPP (quad_perm / row_share), ds_swizzle or readlane, and column j of L
// published through a small per-group LDS buffer.
//
// The reference solves the bordered (n+m)^2 system [C F; F' 0][w; mu] =
// [b; f] by LU with partial pivoting (kriging.rs:527-644, linalg.rs:27-111).
// Here the same system is solved through its Schur complement, with only
// forward substitutions fused into the factorization:
//   yb = L^-1 b,  yz = L^-1 z,  YF = L^-1 F
//   S  = YF' YF,  mu = S^-1 (YF' yb - f)
//   value    = yb'yz - mu' YF'yz            (= w'z)
//   variance = c0 - (yb'yb - mu' YF'yb + mu'f)   (= c0 - w'b0, kriging.rs:625-641)
//   lagrange = mu[0]                         (= w[n])
// Simple kriging (m = 0) and the Vecchia conditional (vecchia.rs:302-334)
// are the m = 0 case of the same routine.
//
// Covariance entries are evaluated by a runtime loop into a lane-private
// LDS staging column and then moved to registers with compile-time
// indices, so the model code appears once per pair of row blocks instead of
// once per entry.
#pragma once

#include <type_traits>

#include "gs_internal.hpp"

namespace gs {

template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// ---- broadcast / reduction inside a group of G lanes --------------------
template <int G, int SRC>
__device__ __forceinline__ int bcast32(int v) {
    if constexpr (G == 1) {
        return v;
    } else if constexpr (G == 2) {
        constexpr int c = SRC | (SRC << 2) | ((2 + SRC) << 4) | ((2 + SRC) << 6);
        return __builtin_amdgcn_mov_dpp(v, c, 0xF, 0xF, false);
    } else if constexpr (G == 4) {
        constexpr int c = SRC | (SRC << 2) | (SRC << 4) | (SRC << 6);
        return __builtin_amdgcn_mov_dpp(v, c, 0xF, 0xF, false);
    } else if constexpr (G == 8) {
        // ds_swizzle bit mode inside 32 lanes: src = (lane & 0x18) | SRC
        return __builtin_amdgcn_ds_swizzle(v, 0x18 | (SRC << 5));
    } else if constexpr (G == 16) {
        return __builtin_amdgcn_mov_dpp(v, 0x150 + SRC, 0xF, 0xF, false);  // row_share
    } else if constexpr (G == 32) {
        // ds_swizzle bit mode: every lane of each 32-lane half reads lane SRC
        return __builtin_amdgcn_ds_swizzle(v, SRC << 5);
    } else {
        static_assert(G == 64, "group size");
        return __builtin_amdgcn_readlane(v, SRC);   // one system per wavefront
    }
}

template <int G, int SRC>
__device__ __forceinline__ double bcast(double v) {
    const int lo = __double2loint(v);
    const int hi = __double2hiint(v);
    return __hiloint2double(bcast32<G, SRC>(hi), bcast32<G, SRC>(lo));
}

template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const int lo = __double2loint(v);
    const int hi = __double2hiint(v);
    return __hiloint2double(__builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, false),
                            __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, false));
}
template <int PAT>
__device__ __forceinline__ double swz_d(double v) {
    const int lo = __double2loint(v);
    const int hi = __double2hiint(v);
    return __hiloint2double(__builtin_amdgcn_ds_swizzle(hi, PAT),
                            __builtin_amdgcn_ds_swizzle(lo, PAT));
}

// Butterfly sum; every lane of the group ends with the bitwise-same value.
template <int G>
__device__ __forceinline__ double gsum(double v) {
    if constexpr (G >= 2) v += dpp_d<0xB1>(v);                  // quad_perm(1,0,3,2)
    if constexpr (G >= 4) v += dpp_d<0x4E>(v);                  // quad_perm(2,3,0,1)
    if constexpr (G >= 8) v += swz_d<0x1F | (4 << 10)>(v);      // lane ^ 4
    if constexpr (G >= 16) v += swz_d<0x1F | (8 << 10)>(v);     // lane ^ 8
    if constexpr (G >= 32) v += swz_d<0x1F | (16 << 10)>(v);    // lane ^ 16
    if constexpr (G >= 64) {
        // the two 32-lane halves: same order in every lane
        const double lo = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 0),
                                           __builtin_amdgcn_readlane(__double2loint(v), 0));
        const double hi = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 32),
                                           __builtin_amdgcn_readlane(__double2loint(v), 32));
        v = lo + hi;
    }
    return v;
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Drift basis (kriging.rs:410-439): OK [1]; UK deg 1 [1, xn]; deg 2 adds
// the monomials xn_a xn_b (a <= b).  M = 0 is simple kriging.
template <int D, int M>
__device__ __forceinline__ void basis(const SolveParams &P, const double *x, double *f) {
    if constexpr (M >= 1) f[0] = 1.0;
    if constexpr (M > 1) {
        double xn[D];
#pragma unroll
        for (int d = 0; d < D; d++) xn[d] = (x[d] - P.center[d]) / P.scale;
#pragma unroll
        for (int d = 0; d < D; d++) f[1 + d] = xn[d];
        if constexpr (M > 1 + D) {
            int idx = 1 + D;
#pragma unroll
            for (int a = 0; a < D; a++)
#pragma unroll
                for (int b = a; b < D; b++) f[idx++] = xn[a] * xn[b];
        }
    }
}

constexpr int kWave = 64;

// exp(-s) for s >= 0 from a 64-entry table: s = k ln2/64 + r with
// |r| <= ln2/128 (Cody-Waite, the fdlibm split of ln2 scaled by 1/64, so
// k * hi is exact), k = 64 m + j, exp(-s) = 2^-m * 2^(-j/64) * e^-r, and
// e^-r - 1 by its degree-5 Taylor polynomial (truncation below 4e-17
// relative on that interval), added to the table entry with one fma.  About
// 1.3 ulp at most, 6 fewer operations than the degree-12 form of exp_neg.
// The table holds 2^(-j/64) rounded to nearest; each kernel that evaluates
// lean covariances copies it to LDS first (load_exp_table).
static __constant__ double kExp2Neg64[64] = {
    0x1.0000000000000p+0, 0x1.fa7c1819e90d8p-1, 0x1.f50765b6e4540p-1, 0x1.efa1bee615a27p-1,
    0x1.ea4afa2a490dap-1, 0x1.e502ee78b3ff6p-1, 0x1.dfc97337b9b5fp-1, 0x1.da9e603db3285p-1,
    0x1.d5818dcfba487p-1, 0x1.d072d4a07897cp-1, 0x1.cb720dcef9069p-1, 0x1.c67f12e57d14bp-1,
    0x1.c199bdd85529cp-1, 0x1.bcc1e904bc1d2p-1, 0x1.b7f76f2fb5e47p-1, 0x1.b33a2b84f15fbp-1,
    0x1.ae89f995ad3adp-1, 0x1.a9e6b5579fdbfp-1, 0x1.a5503b23e255dp-1, 0x1.a0c667b5de565p-1,
    0x1.9c49182a3f090p-1, 0x1.97d829fde4e50p-1, 0x1.93737b0cdc5e5p-1, 0x1.8f1ae99157736p-1,
    0x1.8ace5422aa0dbp-1, 0x1.868d99b4492edp-1, 0x1.82589994cce13p-1, 0x1.7e2f336cf4e62p-1,
    0x1.7a11473eb0187p-1, 0x1.75feb564267c9p-1, 0x1.71f75e8ec5f74p-1, 0x1.6dfb23c651a2fp-1,
    0x1.6a09e667f3bcdp-1, 0x1.6623882552225p-1, 0x1.6247eb03a5585p-1, 0x1.5e76f15ad2148p-1,
    0x1.5ab07dd485429p-1, 0x1.56f4736b527dap-1, 0x1.5342b569d4f82p-1, 0x1.4f9b2769d2ca7p-1,
    0x1.4bfdad5362a27p-1, 0x1.486a2b5c13cd0p-1, 0x1.44e086061892dp-1, 0x1.4160a21f72e2ap-1,
    0x1.3dea64c123422p-1, 0x1.3a7db34e59ff7p-1, 0x1.371a7373aa9cbp-1, 0x1.33c08b26416ffp-1,
    0x1.306fe0a31b715p-1, 0x1.2d285a6e4030bp-1, 0x1.29e9df51fdee1p-1, 0x1.26b4565e27cddp-1,
    0x1.2387a6e756238p-1, 0x1.2063b88628cd6p-1, 0x1.1d4873168b9aap-1, 0x1.1a35beb6fcb75p-1,
    0x1.172b83c7d517bp-1, 0x1.1429aaea92de0p-1, 0x1.11301d0125b51p-1, 0x1.0e3ec32d3d1a2p-1,
    0x1.0b5586cf9890fp-1, 0x1.0874518759bc8p-1, 0x1.059b0d3158574p-1, 0x1.02c9a3e778061p-1,
};
static __shared__ double g_exp_tab[64];

__device__ __forceinline__ void load_exp_table() {
    if (threadIdx.x < 64) g_exp_tab[threadIdx.x] = kExp2Neg64[threadIdx.x];
    __syncthreads();
}

__device__ __forceinline__ double exp_neg_t(double s) {
    s = fmin(s, 800.0);
    const double kf = rint(s * 0x1.71547652b82fep+6);      // 64 / ln2
    double r = fma(-kf, 0x1.62e42fee00000p-7, s);           // ln2/64, high part
    r = fma(-kf, 0x1.a39ef35793c76p-39, r);                 // ln2/64, low part
    const double x = -r;
    double p = 1.0 / 120.0;
    p = fma(p, x, 1.0 / 24.0);
    p = fma(p, x, 1.0 / 6.0);
    p = fma(p, x, 0.5);
    p = fma(p, x, 1.0);
    const double q = p * x;                                 // e^x - 1
    const int k = (int)kf;
    const double t = g_exp_tab[k & 63];
    return ldexp(fma(t, q, t), -(k >> 6));
}

// 1/sqrt(d): hardware estimate plus one Newton step (relative error ~1e-16
// for normal d > 0; d <= 0 gives NaN/inf and is caught by the pivot test).
__device__ __forceinline__ double rsqrt_nr(double d) {
    const double y = __builtin_amdgcn_rsq(d);
    return y * (1.5 - 0.5 * d * y * y);
}

// Per-group LDS column buffer: NB column entries + one slot per lane for
// each of the Q right-hand sides (every lane writes its own slot, readers
// take the owner's: no divergent branch inside the unrolled factorization),
// padded to an odd number of 16-byte units so the groups of one wavefront
// start on different banks.
template <int NB, int Q, int G>
struct ColStride {
    static constexpr int raw = NB + Q * G;
    static constexpr int value = ((raw + 1) / 2 % 2 == 0) ? ((raw + 1) / 2 + 1) * 2 : ((raw + 1) / 2) * 2;
};


// The value, made opaque to the optimizer.  The lane-dependent compares of
// the unrolled factorization then stay where they are used instead of being
// hoisted out of the system loop into long-lived (spilled) lane masks.
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

// Row ownership: in row block r, lane g of a group owns row G*r + g (r
// even) or G*r + G-1-g (r odd).  Blocks r and R-1-r then hold rows summing
// to NB-1 in every lane, so the lower-triangle entries of C split evenly
// over the group.
template <int G>
__device__ __forceinline__ int row_of(int r, int g) { return G * r + ((r & 1) ? G - 1 - g : g); }
template <int G>
constexpr int owner_of(int j) { return ((j / G) & 1) ? G - 1 - j % G : j % G; }

// One lean structure's parameters, loaded once per system into registers.
// Exponential-family kinds share g = 1 - (1 + beta q + gam q^2) exp(-q')
// with q = alpha d and q' = d^2 (Gaussian) or q (model.rs:122-181).
template <int D>
struct LeanRegs {
    double t[D * D];    // anisotropy matrix divided by the range
    double ir2;         // 1 / range^2 (isotropic structures)
    double sill, alpha, beta, gam;
    bool aniso, sph, gau;
    int kind;
};

template <int D>
__device__ __forceinline__ LeanRegs<D> load_lean(const DevStructure &S) {
    LeanRegs<D> L;
    const double ir = S.inv_range;
#pragma unroll
    for (int i = 0; i < D * D; i++) L.t[i] = S.m[i] * ir;
    L.ir2 = ir * ir;
    L.sill = S.sill;
    L.aniso = S.aniso != 0;
    L.kind = S.kind;
    L.sph = S.kind == K_SPH;
    L.gau = S.kind == K_GAU;
    L.alpha = S.kind == K_MAT15 ? 1.7320508075688772 : (S.kind == K_MAT25 ? 2.23606797749979 : 1.0);
    L.beta = (S.kind == K_MAT15 || S.kind == K_MAT25) ? 1.0 : 0.0;
    L.gam = S.kind == K_MAT25 ? 1.0 / 3.0 : 0.0;
    return L;
}

// sqrt(q) for q >= 0: v_rsq on max(q, 1e-300) (so q = 0 gives 0, not
// 0 * inf) and one Markstein correction (about 1 ulp for normal q).
__device__ __forceinline__ double sqrt_lean(double q) {
    const double y = __builtin_amdgcn_rsq(fmax(q, 1e-300));
    const double h = q * y;
    return fma(fma(-h, h, q), 0.5 * y, h);
}

// sill * (1 - g(|T dh| / range)) for a lean structure (ANISO: T is the
// structure's rotation/scaling, else the identity).  Summed over the
// structures this is c0 - gamma(dh) for dh != 0, because c0 = nugget + sum
// sill (model.rs:636-660); the exponential family gives sill p(q) e^-q'
// directly, without the cancellation of 1 - g at long range.
// TAB: exp by the LDS table (the kernel ran load_exp_table); otherwise by
// exp_neg, for kernels that keep their LDS and registers for other work.
// FAM compiles the kind in (1 Matern 3/2, 2 Matern 5/2, 3 exponential,
// 4 Gaussian; 0 reads it from L): the same values, fewer instructions.
template <int D, bool ANISO, bool SPH, bool TAB = true, int FAM = 0>
__device__ __forceinline__ double lean_term(const LeanRegs<D> &L, const double *dh) {
    double h2s = 0.0;
    if constexpr (ANISO) {
#pragma unroll
        for (int k = 0; k < D; k++) {
            double e = 0.0;
#pragma unroll
            for (int d = 0; d < D; d++) e = fma(L.t[k * D + d], dh[d], e);
            h2s = fma(e, e, h2s);
        }
    } else {
#pragma unroll
        for (int d = 0; d < D; d++) h2s = fma(dh[d], dh[d], h2s);
        h2s *= L.ir2;
    }
    const double d = sqrt_lean(h2s);
    double c;
    if constexpr (SPH) {
        c = d >= 1.0 ? 0.0 : 1.0 - (1.5 * d - 0.5 * d * d * d);
    } else {
        double p, x;
        if constexpr (FAM == 1) {
            x = 1.7320508075688772 * d;
            p = 1.0 + x;
        } else if constexpr (FAM == 2) {
            x = 2.23606797749979 * d;
            p = fma(fma(1.0 / 3.0, x, 1.0), x, 1.0);
        } else if constexpr (FAM == 3) {
            x = d;
            p = 1.0;
        } else if constexpr (FAM == 4) {
            x = h2s;
            p = 1.0;
        } else {
            const double q = L.alpha * d;
            p = fma(fma(L.gam, q, L.beta), q, 1.0);
            x = L.gau ? h2s : q;
        }
        c = p * (TAB ? exp_neg_t(x) : exp_neg(x));
    }
    return L.sill * c;
}

// Data-data entries e0 <= e < e1 of one pair of row blocks into the lane's
// LDS staging column (slot e - e0).  Entry e < iA is (row iA, column e);
// entry e >= iA is (row NB-1-iA, column NB-2-e), so every register of the
// pair has a compile-time entry number (with one block, entries past iA
// are unused).  FIRST writes, otherwise the slot accumulates.
template <int D, int NB, bool ANISO, bool SPH, bool FIRST, int FAM = 0>
__device__ __forceinline__ void lean_slots(const LeanRegs<D> &L, const int e0, const int e1,
                                           const int iA, const double (&xA)[D],
                                           const double (&xB)[D], const double *xy,
                                           double *stage, const int lane) {
    // one entry per iteration: unrolled, the compiler interleaves the
    // entries' exp chains and runs out of registers
#pragma unroll 2
    for (int e = e0; e < e1; e++) {
        const bool inA = e < iA;
        const int c = inA ? e : NB - 2 - e;
        double dh[D];
#pragma unroll
        for (int d = 0; d < D; d++) dh[d] = (inA ? xA[d] : xB[d]) - xy[c * D + d];
        const double v = lean_term<D, ANISO, SPH, true, FAM>(L, dh);
        double *slot = stage + (e - e0) * kWave + lane;
        *slot = FIRST ? v : *slot + v;
    }
}

// Staged evaluation of entries e0 <= e < e1 of one pair of row blocks (rows
// iA and NB-1-iA of the lane; one block when R = 1); with TGT the two
// target entries are accumulated into tA / tB as well.  The structure loop
// is outside the entry loop, so each structure's parameters are read once,
// and the kind / anisotropy choice is made once per structure.
template <int D, int NB, bool LEAN, bool TGT>
__device__ __forceinline__ void staged_entries(const DevModel &Mdl, const int e0, const int e1,
                                               const int iA, const double (&xA)[D],
                                               const double (&xB)[D], const double *xy,
                                               const double *tg, double *stage, const int lane,
                                               double &tA, double &tB) {
    double dA[D], dB[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        dA[d] = xA[d] - tg[d];
        dB[d] = xB[d] - tg[d];
    }
    const int ns = Mdl.ns;
    for (int s = 0; s < ns; s++) {
        if constexpr (LEAN) {
            const LeanRegs<D> L = load_lean<D>(Mdl.s[s]);
#define GS_LEAN_CASE(AN, SP, FM)                                                                \
    do {                                                                                        \
        if (s == 0) lean_slots<D, NB, AN, SP, true, FM>(L, e0, e1, iA, xA, xB, xy, stage, lane); \
        else lean_slots<D, NB, AN, SP, false, FM>(L, e0, e1, iA, xA, xB, xy, stage, lane);      \
        if constexpr (TGT) {                                                                    \
            tA += lean_term<D, AN, SP, true, FM>(L, dA);                                        \
            tB += lean_term<D, AN, SP, true, FM>(L, dB);                                        \
        }                                                                                       \
    } while (0)
#define GS_LEAN_KIND(AN)                                                                        \
    do {                                                                                        \
        if (L.kind == K_MAT15) GS_LEAN_CASE(AN, false, 1);                                      \
        else if (L.kind == K_MAT25) GS_LEAN_CASE(AN, false, 2);                                 \
        else if (L.kind == K_EXP) GS_LEAN_CASE(AN, false, 3);                                   \
        else GS_LEAN_CASE(AN, false, 4);                                                        \
    } while (0)
            if (L.aniso) {
                if (L.sph) GS_LEAN_CASE(true, true, 0);
                else GS_LEAN_KIND(true);
            } else {
                if (L.sph) GS_LEAN_CASE(false, true, 0);
                else GS_LEAN_KIND(false);
            }
#undef GS_LEAN_KIND
#undef GS_LEAN_CASE
        } else {
            const DevStructure *S = &Mdl.s[s];
            const bool first = s == 0;
            for (int e = e0; e < e1; e++) {
                const bool inA = e < iA;
                const int c = inA ? e : NB - 2 - e;
                double dh[3] = {0.0, 0.0, 0.0};
#pragma unroll
                for (int d = 0; d < D; d++) dh[d] = (inA ? xA[d] : xB[d]) - xy[c * D + d];
                const double v = general_term<D>(S, Mdl.nodes, dh[0], dh[1], dh[2]);
                double *slot = stage + (e - e0) * kWave + lane;
                *slot = first ? v : *slot + v;
            }
            if constexpr (TGT) {
                tA += general_term<D>(S, Mdl.nodes, dA[0], dA[1], D > 2 ? dA[D - 1] : 0.0);
                tB += general_term<D>(S, Mdl.nodes, dB[0], dB[1], D > 2 ? dB[D - 1] : 0.0);
            }
        }
    }
}

// Covariance entry from the staged structure sum: the lean class sums
// sill (1 - g) (the entry itself), the general class sums sill g.
template <bool LEAN>
__device__ __forceinline__ double entry(double c0, double nug, double s) {
    if constexpr (LEAN) return s;
    else return c0 - (nug + s);
}

// Staging slots per lane: the NB-1 entries of a block pair go through the
// staging column in passes of this many slots (two passes, four for 64-point
// systems, whose larger staging column would otherwise halve the wavefronts
// that fit in LDS).
template <int NB>
constexpr int stage_slots() { return NB == 64 ? 16 : NB / 2; }

// Right-looking Cholesky of the row-distributed C held in a (lane g owns
// row row_of(r, g) of block r, columns before the diagonal), fused with the
// forward substitution of the Q right-hand sides in y.  On return y holds
// L^-1 b, invd[r] = 1/L_ii of the owned row of block r, and a[r][c] for c
// below the row holds L_ic * L_cc (the multipliers before scaling).
// Returns false when a pivot fails the singular test against `thresh`:
// LU (kriging, linalg.rs:42-57) fails when the pivot is below 1e-12 of the
// matrix scale, so GE keeps pivots >= thresh; the Cholesky of the Vecchia
// conditionals (linalg.rs:135-140) fails unless the pivot exceeds 1e-12 of
// the largest diagonal entry (GE = false).
template <int NB, int G, int Q, bool DADD = false, bool GE = true>
__device__ __forceinline__ bool factor_forward(double (&a)[NB / G][NB], double (&y)[Q][NB / G],
                                               const double c0, const double thresh, double *col,
                                               const int g, double (&invd)[NB / G],
                                               const double *dadd = nullptr) {
    constexpr int R = NB / G;
    // dg[r]: the diagonal entry of the owned row of block r, updated as
    // the entry itself would be (d -= l^2 per column).
    // invd[r]: 1/L_ii of the owned row of block r.  The right-hand sides of
    // a row are scaled once at the end instead of being overwritten by the
    // owner at its step.
    int rid[R];
    double dg[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        rid[r] = row_of<G>(r, g);
        if constexpr (DADD) dg[r] = c0 + dadd[r];   // data-data diagonal c0 + err_i
        else dg[r] = c0;
        invd[r] = 0.0;
    }
    bool ok = true;
    static_for<0, NB>([&](auto JJ) {
        constexpr int j = decltype(JJ)::value;
        constexpr int oj = owner_of<G>(j);
        constexpr int rj = j / G;
        const double d = bcast<G, oj>(dg[rj]);
        ok = ok && (GE ? d >= thresh : d > thresh);
        const double inv = rsqrt_nr(d);
        invd[rj] = (g == oj) ? inv : invd[rj];
        double l[R];
#pragma unroll
        for (int r = rj; r < R; r++) l[r] = a[r][j] * inv;
        l[rj] = (rid[rj] > j) ? l[rj] : 0.0;
#pragma unroll
        for (int r = rj; r < R; r++) dg[r] -= l[r] * l[r];
        // Publish column j (rows > j) and the scaled right-hand sides of
        // row j through the group's LDS buffer; every lane then reads the
        // column back with wide broadcast loads.
#pragma unroll
        for (int r = rj; r < R; r++) col[rid[r]] = l[r];
#pragma unroll
        for (int qq = 0; qq < Q; qq++) col[NB + qq * G + g] = y[qq][rj] * inv;
        wave_sync();
#pragma unroll
        for (int qq = 0; qq < Q; qq++) {
            const double yj = col[NB + qq * G + oj];
#pragma unroll
            for (int r = rj; r < R; r++) y[qq][r] -= l[r] * yj;
        }
        static_for<j + 1, NB - 1>([&](auto KK) {
            constexpr int k = decltype(KK)::value;
            const double lk = col[k];
#pragma unroll
            for (int r = k / G; r < R; r++) a[r][k] -= l[r] * lk;
            // at most 8 column values in flight: keeps the scheduler from
            // loading the whole column ahead (register pressure)
            if constexpr (k % 8 == 7) __builtin_amdgcn_sched_barrier(0);
        });
        wave_sync();
    });
#pragma unroll
    for (int qq = 0; qq < Q; qq++)
#pragma unroll
        for (int r = 0; r < R; r++) y[qq][r] *= invd[r];
    return ok;
}


// Core: builds C (rows owned by this lane), factors it and forward-solves
// the Q right-hand sides held in y.  `rowx` are the coordinates of the
// owned rows, `xy` the group's neighbour coordinates in LDS, `tg` the
// target; y[0] is filled here with the covariance to the target.  Entries
// are c0 - gamma(dh) with gamma = nugget + sum sill g (model.rs:636-646);
// gamma(0) = 0, so a target on a data point gets c0.  Returns false when a
// pivot fails the singular test against `thresh` (factor_forward).
//
// With X: ea[r] is added to the diagonal of the owned row of block r
// (measurement error), and when n_off > 0 the target column is the block
// average (1/U) sum_u C(x_i - (tg + off_u)) instead (kriging.rs:709-725).
template <int D, int NB, int G, int Q, bool LEAN, bool X = false, bool GE = true,
          int HS = stage_slots<NB>()>
__device__ __forceinline__ bool build_factor_solve(const DevModel &Mdl, const int cnt,
                                                   const double (&rowx)[NB / G][D],
                                                   const double *xy, const double *tg,
                                                   double *stage, double *col, const int lane,
                                                   const int g_in, const double thresh,
                                                   double (&y)[Q][NB / G],
                                                   const double *ea = nullptr,
                                                   const double *offs = nullptr,
                                                   const int n_off = 0) {
    constexpr int R = NB / G;
    const int g = opaque(g_in);
    const double c0 = Mdl.c0;
    const double nug = Mdl.nugget;
    double a[R][NB];
    constexpr int NE = NB - 1;              // entries per lane and block pair
    constexpr int NPS = (NE + HS - 1) / HS; // staging passes
    static_for<0, (R + 1) / 2>([&](auto PP) {
        constexpr int rA = decltype(PP)::value;
        constexpr int rB = R - 1 - rA;
        constexpr bool single = rA == rB;
        const int iA = row_of<G>(rA, g);
        const int iB = NB - 1 - iA;
        const bool liveA = iA < cnt;
        const bool liveB = iB < cnt;
        double tA = 0.0, tB = 0.0;
        // Entries right of the diagonal receive unused values: the
        // factorization never reads them (its multipliers for rows <= j are
        // selected away) and the diagonal is tracked in dg below.
        static_for<0, NPS>([&](auto PS) {
            constexpr int e0 = decltype(PS)::value * HS;
            constexpr int e1 = e0 + HS < NE ? e0 + HS : NE;
            staged_entries<D, NB, LEAN, decltype(PS)::value == 0>(Mdl, e0, e1, iA, rowx[rA],
                                                                  rowx[rB], xy, tg, stage, lane,
                                                                  tA, tB);
            static_for<0, G * (rA + 1)>([&](auto CC) {
                constexpr int c = decltype(CC)::value;
                if constexpr (c >= e0 && c < e1) {
                    const double sv = stage[(c - e0) * kWave + lane];
                    a[rA][c] = liveA ? entry<LEAN>(c0, nug, sv) : 0.0;
                }
            });
            if constexpr (!single) {
                static_for<0, G * (rB + 1)>([&](auto CC) {
                    constexpr int c = decltype(CC)::value;
                    constexpr int e = NB - 2 - c;
                    if constexpr (e >= e0 && e < e1) {
                        const double sv = stage[(e - e0) * kWave + lane];
                        a[rB][c] = liveB ? entry<LEAN>(c0, nug, sv) : 0.0;
                    }
                });
            }
        });
        if constexpr (G * (rB + 1) == NB) a[rB][NB - 1] = 0.0;   // no entry maps here
        bool zA = true;
#pragma unroll
        for (int d = 0; d < D; d++) zA = zA && rowx[rA][d] == tg[d];
        y[0][rA] = liveA ? (zA ? c0 : entry<LEAN>(c0, nug, tA)) : 0.0;
        if constexpr (!single) {
            bool zB = true;
#pragma unroll
            for (int d = 0; d < D; d++) zB = zB && rowx[rB][d] == tg[d];
            y[0][rB] = liveB ? (zB ? c0 : entry<LEAN>(c0, nug, tB)) : 0.0;
        }
    });

    if constexpr (X) {
        if (n_off > 0) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                double acc = 0.0;
                for (int u = 0; u < n_off; u++) {
                    double dh[D];
#pragma unroll
                    for (int d = 0; d < D; d++) dh[d] = rowx[r][d] - (tg[d] + offs[u * D + d]);
                    acc += cov_offdiag<D, LEAN>(Mdl, dh);
                }
                y[0][r] = row_of<G>(r, g) < cnt ? acc / (double)n_off : 0.0;
            }
        }
    }

    double invd[R];
    if constexpr (X) return factor_forward<NB, G, Q, true, GE>(a, y, c0, thresh, col, g, invd, ea);
    else return factor_forward<NB, G, Q, false, GE>(a, y, c0, thresh, col, g, invd);
}

// Group maximum; every lane of the group ends with the same value.
template <int G>
__device__ __forceinline__ double gmax(double v) {
    if constexpr (G >= 2) v = fmax(v, dpp_d<0xB1>(v));
    if constexpr (G >= 4) v = fmax(v, dpp_d<0x4E>(v));
    if constexpr (G >= 8) v = fmax(v, swz_d<0x1F | (4 << 10)>(v));
    if constexpr (G >= 16) v = fmax(v, swz_d<0x1F | (8 << 10)>(v));
    if constexpr (G >= 32) v = fmax(v, swz_d<0x1F | (16 << 10)>(v));
    if constexpr (G >= 64) {
        const double lo = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 0),
                                           __builtin_amdgcn_readlane(__double2loint(v), 0));
        const double hi = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 32),
                                           __builtin_amdgcn_readlane(__double2loint(v), 32));
        v = fmax(lo, hi);
    }
    return v;
}

// Drift block by modified Gram-Schmidt on the augmented columns
// [YF | yb | yz] (YF = L^-1 F, yb = L^-1 b, yz = L^-1 z, rows distributed
// over the group).  With YF = Q R, g = R^-T f and the residuals
// yb_perp = yb - Q Q'yb, yz_perp = yz - Q Q'yz:
//   value    = yb_perp' yz_perp + g' Q'yz
//   variance = c0 - |yb_perp|^2 - g'(2 Q'yb - g)
//   mu       = R^-1 (Q'yb - g),  lagrange = mu[0]
// This never forms YF'YF, so the drift block's conditioning is not squared.
// A column whose orthogonal remainder falls below 1e-12 of the largest
// drift column norm makes the system singular (rank-deficient drift).
// The reference's LU (linalg.rs:42-57) meets the drift block as the Schur
// complement -S, S = F'C^-1 F = R'R: its partial-pivoting pivots of S must
// reach `thresh` (1e-12 of the bordered matrix's largest entry) as well.
template <int R, int G, int M, int Q>
__device__ __forceinline__ double drift_solve(double (&y)[Q][R], const double (&fT)[M], double c0,
                                              const double thresh, bool &ok, double &var,
                                              double &lag) {
    double rr[M][M], qb[M], qz[M];
    double n0 = 0.0;
    static_for<0, M>([&](auto JJ) {
        constexpr int j = decltype(JJ)::value;
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < R; r++) s += y[2 + j][r] * y[2 + j][r];
        n0 = fmax(n0, gsum<G>(s));
    });
    const double tol = 1e-12 * sqrt(n0);
    static_for<0, M>([&](auto JJ) {
        constexpr int j = decltype(JJ)::value;
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < R; r++) s += y[2 + j][r] * y[2 + j][r];
        const double rjj = sqrt(gsum<G>(s));
        ok = ok && (rjj > tol);
        rr[j][j] = rjj;
        const double inv = 1.0 / rjj;
#pragma unroll
        for (int r = 0; r < R; r++) y[2 + j][r] *= inv;
        static_for<j + 1, M>([&](auto KK) {
            constexpr int k = decltype(KK)::value;
            double t = 0.0;
#pragma unroll
            for (int r = 0; r < R; r++) t += y[2 + j][r] * y[2 + k][r];
            t = gsum<G>(t);
            rr[j][k] = t;
#pragma unroll
            for (int r = 0; r < R; r++) y[2 + k][r] -= t * y[2 + j][r];
        });
        double tb = 0.0, tz = 0.0;
#pragma unroll
        for (int r = 0; r < R; r++) {
            tb += y[2 + j][r] * y[0][r];
            tz += y[2 + j][r] * y[1][r];
        }
        tb = gsum<G>(tb);
        tz = gsum<G>(tz);
        qb[j] = tb;
        qz[j] = tz;
#pragma unroll
        for (int r = 0; r < R; r++) {
            y[0][r] -= tb * y[2 + j][r];
            y[1][r] -= tz * y[2 + j][r];
        }
    });
    // Partial pivoting keeps every |u_kk| <= 2^M max|S| (the growth bound),
    // and prod |u_kk| = det S = prod r_jj^2, so every pivot is at least
    // det S / (2^M max|S|)^(M-1); max|S| is the largest diagonal entry (S is
    // positive definite).  Only a system below that bound (near-singular
    // drift) needs the exact pivots, so the elimination below is skipped
    // for all others.
    double det = 1.0, smax = 0.0;
#pragma unroll
    for (int j = 0; j < M; j++) {
        det *= rr[j][j] * rr[j][j];
        double dj = 0.0;
#pragma unroll
        for (int k = 0; k <= j; k++) dj += rr[k][j] * rr[k][j];
        smax = fmax(smax, dj);
    }
    double growth = 1.0;
#pragma unroll
    for (int j = 1; j < M; j++) growth *= (double)(1 << M) * smax;
    if (M == 1) {
        ok = ok && !(det < thresh);
    } else if (!(det >= 2.0 * thresh * growth)) {
        double S[M][M];
#pragma unroll
        for (int i = 0; i < M; i++)
#pragma unroll
            for (int j = 0; j < M; j++) {
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k <= (i < j ? i : j); k++) acc += rr[k][i] * rr[k][j];
                S[i][j] = acc;
            }
        static_for<0, M>([&](auto KK) {
            constexpr int k = decltype(KK)::value;
            double mx = fabs(S[k][k]);
            int p = k;
#pragma unroll
            for (int i = k + 1; i < M; i++) {
                const double v = fabs(S[i][k]);
                p = v > mx ? i : p;
                mx = v > mx ? v : mx;
            }
            ok = ok && !(mx < thresh);
#pragma unroll
            for (int i = k + 1; i < M; i++) {
                const bool sw = p == i;
#pragma unroll
                for (int j = k; j < M; j++) {
                    const double t = S[k][j];
                    S[k][j] = sw ? S[i][j] : t;
                    S[i][j] = sw ? t : S[i][j];
                }
            }
#pragma unroll
            for (int i = k + 1; i < M; i++) {
                const double f = S[i][k] / S[k][k];
#pragma unroll
                for (int j = k + 1; j < M; j++) S[i][j] -= f * S[k][j];
            }
        });
    }
    double pbz = 0.0, pbb = 0.0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        pbz += y[0][r] * y[1][r];
        pbb += y[0][r] * y[0][r];
    }
    pbz = gsum<G>(pbz);
    pbb = gsum<G>(pbb);
    double gg[M];
#pragma unroll
    for (int j = 0; j < M; j++) {        // R' g = f
        double s = fT[j];
#pragma unroll
        for (int i = 0; i < j; i++) s -= rr[i][j] * gg[i];
        gg[j] = s / rr[j][j];
    }
    double mu[M];
#pragma unroll
    for (int j = M - 1; j >= 0; j--) {   // R mu = Q'yb - g
        double s = qb[j] - gg[j];
#pragma unroll
        for (int k = j + 1; k < M; k++) s -= rr[j][k] * mu[k];
        mu[j] = s / rr[j][j];
    }
    double value = pbz, red = pbb;
#pragma unroll
    for (int j = 0; j < M; j++) {
        value += gg[j] * qz[j];
        red += gg[j] * (2.0 * qb[j] - gg[j]);
    }
    var = c0 - red;
    lag = mu[0];
    return value;
}

// Wavefronts per SIMD the lean kernels are compiled for (the occupancy
// target handed to the register allocator).  At NB = 32,
// G = 16, M <= 1: three (168 VGPRs, no spills; registers and LDS both allow
// three).  At NB = 64, G = 32: two, which costs 30-58 VGPRs spilled to
// scratch but doubles the wavefronts that hide the factorization's LDS
// round trips (the 3-D, M = 4 solve of
// 4 * 10^6 octant systems: 108 -> 77 ms).  The general-class and extended
// instantiations keep the compiler's choice: they would spill far more.
template <int NB, int G, int M, bool LEAN, bool X>
constexpr int krige_waves() {
    if (NB == 32 && G == 16 && M <= 1) return 3;
    if (NB == 64 && G == 32 && LEAN && !X) return 2;
    return 1;
}

// The 2-D, 32-point Vecchia kernel at 16 lanes per point fits three
// wavefronts per SIMD (<= 168 VGPRs, LDS below a twelfth of the CU's).
template <int D, int NB, int G, bool LEAN>
constexpr int vecchia_waves() { return (D == 2 && NB == 32 && G == 16 && LEAN) ? 3 : 1; }
// Its staging passes are 10 slots wide (four passes over the 31 entries of
// a lane) so that the LDS of three wavefronts per SIMD fits.
template <int D, int NB, int G, bool LEAN>
constexpr int vecchia_stage() {
    return vecchia_waves<D, NB, G, LEAN>() == 3 ? 10 : stage_slots<NB>();
}

// X = true adds the extended systems: measurement error (kriging.rs:181-218,
// 544), external drift (223-260, 431-437) and block kriging (678-777).
template <int D, int NB, int G, int M, bool LEAN, bool X>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(krige_waves<NB, G, M, LEAN, X>()))) void krige_fp64_kernel(SolveParams P) {
    constexpr int R = NB / G;
    constexpr int Q = 2 + M;
    constexpr int GPW = kWave / G;
    __shared__ double s_xy[GPW * NB * D];
    __shared__ double s_stage[stage_slots<NB>() * kWave];
    __shared__ __attribute__((aligned(16))) double s_col[GPW * ColStride<NB, Q, G>::value];
    const int lane = threadIdx.x;
    const int g = lane % G;
    const int grp = lane / G;
    double *xy = s_xy + grp * NB * D;
    double *colbuf = s_col + grp * ColStride<NB, Q, G>::value;
    const DevModel &Mdl = *P.model;
    const double c0 = Mdl.c0;
    const int64_t nsys = P.list ? (int64_t)(*P.list_len) : P.T;
    if constexpr (LEAN) load_exp_table();

    for (int64_t sys = xcd_block(blockIdx.x, gridDim.x) * GPW + grp; sys < nsys;
         sys += (int64_t)gridDim.x * GPW) {
        const int64_t t = P.list ? (int64_t)P.list[sys] : sys;
        const int cnt = P.count[t];
        if (cnt < P.min_nb || cnt < M || cnt > NB) {
            if (g == 0) {
                P.value[t] = __builtin_nan("");
                P.variance[t] = __builtin_nan("");
                if (P.lagrange) P.lagrange[t] = __builtin_nan("");
                if (P.status)
                    P.status[t] = cnt > NB ? 4 : (cnt < P.min_nb ? 5 : 1);
            }
            continue;
        }
        double tg[D];
#pragma unroll
        for (int d = 0; d < D; d++) tg[d] = P.targets[t * D + d];
        const int32_t *nbl = P.nbr + t * (int64_t)P.stride;
        // entries at or right of the diagonal are never read (their
        // multipliers are selected away), so slots past the count stay as
        // the previous system left them
        for (int j = g; j < cnt; j += G) {
            const int64_t id = nbl[j];
#pragma unroll
            for (int d = 0; d < D; d++) xy[j * D + d] = P.coords[id * D + d];
        }
        double zr[R];
        double ea[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = row_of<G>(r, g);
            zr[r] = i < cnt ? P.values[nbl[i]] : 0.0;
            if constexpr (X) ea[r] = (P.err && i < cnt) ? P.err[nbl[i]] : 0.0;
        }
        wave_sync();
        double rowx[R][D];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = row_of<G>(r, g);
#pragma unroll
            for (int d = 0; d < D; d++) rowx[r][d] = i < cnt ? xy[i * D + d] : 0.0;
        }
        double y[Q][R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = row_of<G>(r, g);
            const bool live = i < cnt;
            y[1][r] = live ? (M == 0 ? zr[r] - P.sk_mean : zr[r]) : 0.0;
            if constexpr (M > 0) {
                double f[M];
                if constexpr (X && M > 1) {
                    if (P.xdrift) {
                        f[0] = 1.0;
#pragma unroll
                        for (int q = 1; q < M; q++)
                            f[q] = live ? P.xdrift[(int64_t)nbl[i] * (M - 1) + (q - 1)] : 0.0;
                    } else {
                        basis<D, M>(P, rowx[r], f);
                    }
                } else {
                    basis<D, M>(P, rowx[r], f);
                }
#pragma unroll
                for (int q = 0; q < M; q++) y[2 + q][r] = live ? f[q] : 0.0;
            }
        }
        // Scale of the reference's singular test (linalg.rs:36-39): the
        // largest |entry| of the bordered matrix, max(c0 + err_i, |F_iq|)
        // over the live rows (covariances never exceed c0 in magnitude).
        // Without the extended systems that is c0 for simple kriging and
        // max(c0, 1) with a drift: the constant column is 1 and the
        // polynomial basis of a datum stays within [-1, 1], because the
        // basis is centred and scaled on the data's bounding box
        // (kriging.rs:340-347).  Uniform over the wavefront, so it stays in
        // scalar registers through the factorization.
        double amax = M >= 1 ? fmax(c0, 1.0) : c0;
        if constexpr (X) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                if (row_of<G>(r, g) < cnt) {
                    amax = fmax(amax, c0 + ea[r]);
#pragma unroll
                    for (int q = 1; q < M; q++) amax = fmax(amax, fabs(y[2 + q][r]));
                }
            }
            amax = gmax<G>(amax);
        }
        const double thresh = 1e-12 * amax;
        bool ok;
        if constexpr (X)
            ok = build_factor_solve<D, NB, G, Q, LEAN, true>(Mdl, cnt, rowx, xy, tg, s_stage, colbuf,
                                                             lane, g, thresh, y, ea, P.offsets,
                                                             P.n_off);
        else
            ok = build_factor_solve<D, NB, G, Q, LEAN>(Mdl, cnt, rowx, xy, tg, s_stage, colbuf, lane,
                                                       g, thresh, y);
        wave_sync();  // the group's LDS rows are rewritten by its next system
        // variance base: c0, or the within-block covariance for a block
        double vbase = c0;
        if constexpr (X) vbase = P.offsets ? *P.cbb : c0;

        double value, var, lag;
        if constexpr (M == 0) {
            double bz = 0.0, bb = 0.0;
#pragma unroll
            for (int r = 0; r < R; r++) {
                bz += y[0][r] * y[1][r];
                bb += y[0][r] * y[0][r];
            }
            value = P.sk_mean + gsum<G>(bz);
            var = vbase - gsum<G>(bb);
            lag = __builtin_nan("");
        } else {
            double fT[M];
            if constexpr (X && M > 1) {
                if (P.xdrift) {
                    fT[0] = 1.0;
#pragma unroll
                    for (int q = 1; q < M; q++)
                        fT[q] = (P.tdrift[t * (M - 1) + (q - 1)] - P.xmean[q - 1]) / P.xscale[q - 1];
                } else {
                    basis<D, M>(P, tg, fT);
                }
            } else {
                basis<D, M>(P, tg, fT);
            }
            value = drift_solve<R, G, M>(y, fT, vbase, thresh, ok, var, lag);
        }
        if (g == 0) {
            if (ok) {
                P.value[t] = value;
                P.variance[t] = fmax(var, 0.0);
                if (P.lagrange) P.lagrange[t] = lag;
                if (P.status) P.status[t] = 0;
            } else {
                P.value[t] = __builtin_nan("");
                P.variance[t] = __builtin_nan("");
                if (P.lagrange) P.lagrange[t] = __builtin_nan("");
                if (P.status) P.status[t] = 3;
            }
        }
    }
}

// ---- Vecchia log-likelihood ---------------------------------------------
// loglik_with_plan (vecchia.rs:285-336) for many models at once.  A group
// handles one ordered point and loops over the models, reusing the
// gathered conditioning coordinates.  Per-point terms are summed in a
// fixed order (group-serial, then groups in index order inside the block,
// then blocks in index order by vecchia_reduce), so the result does not
// depend on scheduling.

template <int D, int NB, int G, bool LEAN>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(vecchia_waves<D, NB, G, LEAN>()))) void vecchia_kernel(VecchiaParams P) {
    constexpr int R = NB / G;
    constexpr int GPW = kWave / G;
    constexpr int HS = vecchia_stage<D, NB, G, LEAN>();
    __shared__ double s_xy[GPW * NB * D];
    __shared__ double s_z[GPW * NB];
    __shared__ double s_stage[HS * kWave];
    __shared__ double s_sum[GPW * kMaxModelsPerLaunch];
    __shared__ __attribute__((aligned(16))) double s_col[GPW * ColStride<NB, 2, G>::value];
    const int lane = threadIdx.x;
    const int g = lane % G;
    const int grp = lane / G;
    double *xy = s_xy + grp * NB * D;
    double *zs = s_z + grp * NB;
    double *colbuf = s_col + grp * ColStride<NB, 2, G>::value;
    const int nm = P.n_models;
    for (int p = g; p < nm; p += G) s_sum[grp * kMaxModelsPerLaunch + p] = 0.0;
    if constexpr (LEAN) load_exp_table();
    wave_sync();
    const double ln2pi = 1.8378770664093453;  // ln(2 pi)

    const int64_t k0 = (int64_t)blockIdx.x * P.points_per_block;
    int64_t k1 = k0 + P.points_per_block;
    if (k1 > P.n) k1 = P.n;
    for (int64_t k = k0 + grp; k < k1; k += GPW) {
        const int64_t self = P.order[k];
        const int cnt = P.nb_count[k];
        const int32_t *nbl = P.nb + k * (int64_t)P.m;
        double tg[D];
#pragma unroll
        for (int d = 0; d < D; d++) tg[d] = P.coords[self * D + d];
        const double zi = P.zc[self];
        // slots past the count are zero: every entry a kernel evaluates,
        // used or not, then comes from this system's own finite data
        for (int j = g; j < NB; j += G) {
            const int64_t id = j < cnt ? nbl[j] : 0;
#pragma unroll
            for (int d = 0; d < D; d++) xy[j * D + d] = j < cnt ? P.coords[id * D + d] : 0.0;
            zs[j] = j < cnt ? P.zc[id] : 0.0;
        }
        wave_sync();
        for (int p = 0; p < nm; p++) {
            const DevModel &Mdl = P.models[p];
            const double sill = Mdl.c0;
            double mu = 0.0, var = sill;
            bool ok = true;
            if (cnt > 0) {
                // the owned rows' coordinates and values, re-read from LDS
                // for every model rather than held across the model loop
                double y[2][R], rowx[R][D];
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int i = row_of<G>(r, g);
                    y[1][r] = zs[i];
#pragma unroll
                    for (int d = 0; d < D; d++) rowx[r][d] = xy[i * D + d];
                }
                ok = build_factor_solve<D, NB, G, 2, LEAN, false, false, HS>(
                    Mdl, cnt, rowx, xy, tg, s_stage, colbuf, lane, g, 1e-12 * sill, y);
                double s01 = 0.0, s00 = 0.0;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    s01 += y[0][r] * y[1][r];
                    s00 += y[0][r] * y[0][r];
                }
                mu = gsum<G>(s01);
                const double red = gsum<G>(s00);
                var = fmax(sill - red, 1e-12);
            }
            const double rr = zi - mu;
            const double ll = -0.5 * (ln2pi + log(var) + rr * rr / var);
            if (g == 0) {
                s_sum[grp * kMaxModelsPerLaunch + p] += ll;
                if (!ok) atomicMax(&P.status[p], 3);
            }
        }
        wave_sync();
    }
    __syncthreads();
    for (int p = lane; p < nm; p += kWave) {
        double s = 0.0;
        for (int q = 0; q < GPW; q++) s += s_sum[q * kMaxModelsPerLaunch + p];
        P.partial[(int64_t)blockIdx.x * nm + p] = s;
    }
}

// ---- Vecchia log-likelihood gradient --------------------------------------
// loglik_grad_inner (vecchia.rs:428-566) for many single-structure isotropic
// models of the five closed-form kinds (g_and_dg, vecchia.rs:353-385).
// Per ordered point, with d = h / range, correlations r_a = 1 - g(d_ia),
// R_ab = 1 - g(d_ab) (diagonal 1), K = nugget I + psill R, k = psill r, and
// the range-derivative factors rho_a = d_ia g'(d_ia), P_ab = d_ab g'(d_ab):
// factor K = L L' once and solve yr = L^-1 r, yz = L^-1 z,
// wr = L^-T yr (= K^-1 r), v = L^-T yz (= K^-1 z).  Then with w = psill wr
// (the reference's K^-1 k):
//   mu = psill yr.yz, reduction = psill^2 |yr|^2, var = max(sill - reduction, 1e-12)
//   d/dnugget: dmu = -psill wr.v,  dvar = 1 + psill^2 |wr|^2
//   d/dpsill:  dmu = nugget wr.v,  dvar = 1 - psill (nugget |wr|^2 + |yr|^2)
//   d/drange:  dmu = psill/range (rho.v - psill wr'Pv),
//              dvar = -psill^2/range (2 rho.wr - psill wr'P wr)
// which are the reference's solves with right-hand sides dk/dtheta -
// dK/dtheta w written through the same factor (dK/dnugget = I, dK/dpsill =
// R, dK/drange = psill/range P); every dvar is zeroed where the variance is
// clamped, as the reference does.  The per-point terms
// -0.5 dvar/var + r dmu/var + 0.5 r^2 dvar/var^2 (r = z_i - mu) and the
// log-density are summed per model in the same fixed order as the
// log-likelihood kernel.  Solving with r instead of k keeps psill = 0 exact.

template <int D>
struct GradRegs {
    LeanRegs<D> L;
    double r1, r2, r3;   // d g'(d) = d (r1 + d (r2 + d r3)) e^-q' (exponential family)
};

template <int D>
__device__ __forceinline__ GradRegs<D> load_grad(const DevStructure &S) {
    GradRegs<D> R;
    R.L = load_lean<D>(S);
    R.r1 = S.kind == K_EXP ? 1.0 : 0.0;
    R.r2 = S.kind == K_GAU ? 2.0 : (S.kind == K_MAT15 ? 3.0 : (S.kind == K_MAT25 ? 5.0 / 3.0 : 0.0));
    R.r3 = S.kind == K_MAT25 ? 5.0 * 2.23606797749979 / 3.0 : 0.0;
    return R;
}

// Correlation c = 1 - g(d) and rho = d g'(d) for an isotropic separation.
template <int D, bool SPH>
__device__ __forceinline__ void grad_pair(const GradRegs<D> &R, const double *dh, double &c,
                                          double &rho) {
    double h2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; d++) h2 = fma(dh[d], dh[d], h2);
    const double h2s = h2 * R.L.ir2;
    const double d = sqrt_lean(h2s);
    if constexpr (SPH) {
        const bool in = d < 1.0;
        c = in ? 1.0 - (1.5 * d - 0.5 * d * d * d) : 0.0;
        rho = in ? d * (1.5 - 1.5 * d * d) : 0.0;
    } else {
        const double q = R.L.alpha * d;
        const double e = exp_neg_t(R.L.gau ? h2s : q);
        c = fma(fma(R.L.gam, q, R.L.beta), q, 1.0) * e;
        rho = d * fma(fma(R.r3, d, R.r2), d, R.r1) * e;
    }
}

// Entries e0 <= e < e1 of block pair bp: correlations to the staging column,
// P to the lane's persistent column (zero for rows past the count).
template <int D, int NB, bool SPH>
__device__ __forceinline__ void grad_slots(const GradRegs<D> &R, const int e0, const int e1,
                                           const int iA, const bool liveA, const bool liveB,
                                           const double (&xA)[D], const double (&xB)[D],
                                           const double *xy, double *stage, double *pcol,
                                           const int lane) {
#pragma unroll 1
    for (int e = e0; e < e1; e++) {
        const bool inA = e < iA;
        const int c = inA ? e : NB - 2 - e;
        double dh[D];
#pragma unroll
        for (int d = 0; d < D; d++) dh[d] = (inA ? xA[d] : xB[d]) - xy[c * D + d];
        double cv, rho;
        grad_pair<D, SPH>(R, dh, cv, rho);
        stage[(e - e0) * kWave + lane] = cv;
        pcol[e * kWave + lane] = (inA ? liveA : liveB) ? rho : 0.0;
    }
}

template <int D, int NB, int G>
__global__ __launch_bounds__(kWave) void vecchia_grad_kernel(VecchiaParams P) {
    constexpr int R = NB / G;
    static_assert(R >= 2 && R % 2 == 0, "gradient kernel pairs row blocks");
    constexpr int GPW = kWave / G;
    constexpr int NE = NB - 1;
    constexpr int HS = stage_slots<NB>();
    constexpr int NP = R / 2;
    constexpr int kQ = 4;                       // loglik, d/dnugget, d/dpsill, d/drange
    __shared__ double s_xy[GPW * NB * D];
    __shared__ double s_stage[HS * kWave];
    __shared__ double s_p[NP * NE * kWave];
    __shared__ double s_wv[GPW * 2 * NB];
    __shared__ double s_sum[GPW * kMaxModelsPerLaunch * kQ];
    __shared__ __attribute__((aligned(16))) double s_col[GPW * ColStride<NB, 2, G>::value];
    const int lane = threadIdx.x;
    const int g = lane % G;
    const int grp = lane / G;
    double *xy = s_xy + grp * NB * D;
    double *colbuf = s_col + grp * ColStride<NB, 2, G>::value;
    double *wv = s_wv + grp * 2 * NB;
    double *sums = s_sum + grp * kMaxModelsPerLaunch * kQ;
    const int nm = P.n_models;
    for (int q = g; q < nm * kQ; q += G) sums[q] = 0.0;
    load_exp_table();
    wave_sync();
    const double ln2pi = 1.8378770664093453;   // ln(2 pi)

    const int64_t k0 = (int64_t)blockIdx.x * P.points_per_block;
    int64_t k1 = k0 + P.points_per_block;
    if (k1 > P.n) k1 = P.n;
    for (int64_t k = k0 + grp; k < k1; k += GPW) {
        const int64_t self = P.order[k];
        const int cnt = P.nb_count[k];
        const int32_t *nbl = P.nb + k * (int64_t)P.m;
        double tg[D];
#pragma unroll
        for (int d = 0; d < D; d++) tg[d] = P.coords[self * D + d];
        const double zi = P.zc[self];
        // slots past the count are zero: every entry a kernel evaluates,
        // used or not, then comes from this system's own finite data
        for (int j = g; j < NB; j += G) {
            const int64_t id = j < cnt ? nbl[j] : 0;
#pragma unroll
            for (int d = 0; d < D; d++) xy[j * D + d] = j < cnt ? P.coords[id * D + d] : 0.0;
        }
        double zr[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = row_of<G>(r, g);
            zr[r] = i < cnt ? P.zc[nbl[i]] : 0.0;
        }
        wave_sync();
        double rowx[R][D];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = row_of<G>(r, g);
#pragma unroll
            for (int d = 0; d < D; d++) rowx[r][d] = i < cnt ? xy[i * D + d] : 0.0;
        }
        for (int p = 0; p < nm; p++) {
            const DevModel &Mdl = P.models[p];
            const double nug = Mdl.nugget;
            const double sill = Mdl.c0;
            const double psill = Mdl.s[0].sill;
            const double ir = Mdl.s[0].inv_range;
            double ll, gn, gp, gr;
            bool ok = true;
            if (cnt == 0) {
                // first point: N(0, sill); only nugget and psill move sill
                const double var = sill, r = zi;
                ll = -0.5 * (ln2pi + log(var) + r * r / var);
                const double common = -0.5 / var + 0.5 * r * r / (var * var);
                gn = common;
                gp = common;
                gr = 0.0;
            } else {
                const int gg = opaque(g);
                const GradRegs<D> GR = load_grad<D>(Mdl.s[0]);
                const bool sph = GR.L.sph;
                double a[R][NB], y[2][R], rho[R];
                static_for<0, NP>([&](auto BP) {
                    constexpr int bp = decltype(BP)::value;
                    constexpr int rA = bp, rB = R - 1 - bp;
                    const int iA = row_of<G>(rA, gg);
                    const int iB = NB - 1 - iA;
                    const bool liveA = iA < cnt, liveB = iB < cnt;
                    double *pcol = s_p + (size_t)bp * NE * kWave;
                    static_for<0, (NE + HS - 1) / HS>([&](auto PS) {
                        constexpr int e0 = decltype(PS)::value * HS;
                        constexpr int e1 = e0 + HS < NE ? e0 + HS : NE;
                        if (sph) grad_slots<D, NB, true>(GR, e0, e1, iA, liveA, liveB, rowx[rA], rowx[rB], xy, s_stage, pcol, lane);
                        else grad_slots<D, NB, false>(GR, e0, e1, iA, liveA, liveB, rowx[rA], rowx[rB], xy, s_stage, pcol, lane);
                        static_for<0, G * (rA + 1)>([&](auto CC) {
                            constexpr int c = decltype(CC)::value;
                            if constexpr (c >= e0 && c < e1)
                                a[rA][c] = liveA ? psill * s_stage[(c - e0) * kWave + lane] : 0.0;
                        });
                        static_for<0, G * (rB + 1)>([&](auto CC) {
                            constexpr int c = decltype(CC)::value;
                            constexpr int e = NB - 2 - c;
                            if constexpr (e >= e0 && e < e1)
                                a[rB][c] = liveB ? psill * s_stage[(e - e0) * kWave + lane] : 0.0;
                        });
                    });
                    if constexpr (G * (rB + 1) == NB) a[rB][NB - 1] = 0.0;
                });
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const bool live = row_of<G>(r, gg) < cnt;
                    double dh[D];
#pragma unroll
                    for (int d = 0; d < D; d++) dh[d] = rowx[r][d] - tg[d];
                    double cv, rv;
                    if (sph) grad_pair<D, true>(GR, dh, cv, rv);
                    else grad_pair<D, false>(GR, dh, cv, rv);
                    y[0][r] = live ? cv : 0.0;
                    y[1][r] = zr[r];
                    rho[r] = live ? rv : 0.0;
                }
                double invd[R];
                ok = factor_forward<NB, G, 2, false, false>(a, y, sill, 1e-12 * sill, colbuf, gg, invd);
                // back substitution L' x = y for both right-hand sides; a lane
                // only needs x of its own rows
                double x[2][R];
#pragma unroll
                for (int r = 0; r < R; r++) x[0][r] = x[1][r] = 0.0;
                int rid[R];
#pragma unroll
                for (int r = 0; r < R; r++) rid[r] = row_of<G>(r, gg);
                static_for<0, NB>([&](auto JJ) {
                    constexpr int j = NB - 1 - decltype(JJ)::value;
                    constexpr int oj = owner_of<G>(j);
                    constexpr int rj = j / G;
                    double s0 = 0.0, s1 = 0.0;
#pragma unroll
                    for (int r = rj; r < R; r++) {
                        const bool below = rid[r] > j;
                        s0 += below ? a[r][j] * x[0][r] : 0.0;
                        s1 += below ? a[r][j] * x[1][r] : 0.0;
                    }
                    s0 = gsum<G>(s0);
                    s1 = gsum<G>(s1);
                    const bool own = gg == oj;
                    x[0][rj] = own ? (y[0][rj] - invd[rj] * s0) * invd[rj] : x[0][rj];
                    x[1][rj] = own ? (y[1][rj] - invd[rj] * s1) * invd[rj] : x[1][rj];
                });
#pragma unroll
                for (int r = 0; r < R; r++) {
                    wv[rid[r]] = x[0][r];
                    wv[NB + rid[r]] = x[1][r];
                }
                wave_sync();
                double yr2 = 0.0, yrz = 0.0, w2 = 0.0, wvv = 0.0, rw = 0.0, rvv = 0.0;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    yr2 += y[0][r] * y[0][r];
                    yrz += y[0][r] * y[1][r];
                    w2 += x[0][r] * x[0][r];
                    wvv += x[0][r] * x[1][r];
                    rw += rho[r] * x[0][r];
                    rvv += rho[r] * x[1][r];
                }
                // bilinear forms wr'P wr and wr'P v over the lane's entries
                double b1 = 0.0, b2 = 0.0;
                static_for<0, NP>([&](auto BP) {
                    constexpr int bp = decltype(BP)::value;
                    constexpr int rA = bp, rB = R - 1 - bp;
                    const int iA = rid[rA];
                    const double *pcol = s_p + (size_t)bp * NE * kWave;
                    for (int e = 0; e < NE; e++) {
                        const bool inA = e < iA;
                        const int c = inA ? e : NB - 2 - e;
                        const double pe = pcol[e * kWave + lane];
                        const double wa = inA ? x[0][rA] : x[0][rB];
                        const double va = inA ? x[1][rA] : x[1][rB];
                        const double wb = wv[c], vb = wv[NB + c];
                        b1 += pe * (wa * wb);
                        b2 += pe * fma(wa, vb, wb * va);
                    }
                });
                yr2 = gsum<G>(yr2);
                yrz = gsum<G>(yrz);
                w2 = gsum<G>(w2);
                wvv = gsum<G>(wvv);
                rw = gsum<G>(rw);
                rvv = gsum<G>(rvv);
                b1 = 2.0 * gsum<G>(b1);
                b2 = gsum<G>(b2);
                wave_sync();   // wv and the staging columns are rewritten next
                const double mu = psill * yrz;
                const double red = psill * psill * yr2;
                const double raw = sill - red;
                const double var = fmax(raw, 1e-12);
                const bool clamped = raw <= 1e-12;
                const double r = zi - mu;
                ll = -0.5 * (ln2pi + log(var) + r * r / var);
                const double inv_var = 1.0 / var;
                auto term = [&](double dvar, double dmu) {
                    dvar = clamped ? 0.0 : dvar;
                    return -0.5 * dvar * inv_var + r * dmu * inv_var + 0.5 * r * r * dvar * inv_var * inv_var;
                };
                gn = term(1.0 + psill * psill * w2, -psill * wvv);
                gp = term(1.0 - psill * (nug * w2 + yr2), nug * wvv);
                gr = term(-psill * psill * ir * (2.0 * rw - psill * b1), psill * ir * (rvv - psill * b2));
            }
            if (g == 0) {
                sums[p * kQ + 0] += ll;
                sums[p * kQ + 1] += gn;
                sums[p * kQ + 2] += gp;
                sums[p * kQ + 3] += gr;
                if (!ok) atomicMax(&P.status[p], 3);
            }
        }
        wave_sync();
    }
    __syncthreads();
    for (int c = lane; c < nm * kQ; c += kWave) {
        double acc = 0.0;
        for (int q = 0; q < GPW; q++) acc += s_sum[q * kMaxModelsPerLaunch * kQ + c];
        P.partial[(int64_t)blockIdx.x * nm * kQ + c] = acc;
    }
}

template <int D, int NB, int G>
hipError_t launch_vecchia_grad_inst(const VecchiaParams &p, int64_t n_blocks, hipStream_t s) {
    hipLaunchKernelGGL((vecchia_grad_kernel<D, NB, G>), dim3((unsigned)n_blocks), dim3(kWave), 0, s, p);
    return hipGetLastError();
}

// ---- launchers (instantiated per template argument set) -------------------
template <int D, int NB, int M, int G, bool LEAN, bool X>
hipError_t launch_krige_inst(const SolveParams &p, int64_t grid_groups, hipStream_t s) {
    constexpr int GPW = kWave / G;
    int64_t blocks = (grid_groups + GPW - 1) / GPW;
    if (blocks < 1) blocks = 1;
    if (blocks > (1 << 30)) blocks = 1 << 30;
    hipLaunchKernelGGL((krige_fp64_kernel<D, NB, G, M, LEAN, X>), dim3((unsigned)blocks),
                       dim3(kWave), 0, s, p);
    return hipGetLastError();
}

template <int D, int NB, int G, bool LEAN>
hipError_t launch_vecchia_inst(const VecchiaParams &p, int64_t n_blocks, hipStream_t s) {
    hipLaunchKernelGGL((vecchia_kernel<D, NB, G, LEAN>), dim3((unsigned)n_blocks), dim3(kWave), 0, s, p);
    return hipGetLastError();
}

}  // namespace gs
