// FP32-refined kriging and Vecchia solves on gfx950 (CDNA4).
//
// The same group-per-system layout as the FP64 kernels (gs_solve_impl.hpp):
// lane g of a group of G owns rows row_of(r, g) of the system.  The bulk of
// the work runs in FP32:
//   1. covariance entries from FP64 separations (the difference of two
//      coordinates is taken in FP64, then rounded), evaluated in FP32 with
//      the hardware exp / sqrt;
//   2. right-looking Cholesky A32 = L L' with forward substitution of
//      [b | z | F], the drift block in the forward space, and back
//      substitution of the three solution vectors
//        w0 ~ A^-1 (b - F mu0),  v ~ A^-1 (z - F vmu),  p ~ A^-1 F / S;
//   3. one FP64 pass that re-evaluates every covariance entry of the
//      system in FP64 and corrects the outputs with the residual of the
//      FP64 bordered system K = [A F; F' 0] (an adjoint, or dual-weighted,
//      correction): with u0 = [w0; mu0], rhs = [b; f] and r = rhs - K u0,
//        value    = z'w0 + [v; vmu]' r
//        variance = c0 - (2 u0'rhs - u0'K u0)
//        lagrange = mu0 + [p; pmu]' r
//      Each formula is exact up to a term quadratic in the FP32 solve error
//      (for value: ([v; vmu] - K^-1 [z; 0])' K (u0 - K^-1 rhs)), so one pass
//      takes an FP32 error of cond * 6e-8 to (cond * 6e-8)^2.  The residual
//      forms are sums over the entries of A, which is symmetric, so every
//      lane sums over its own lower-triangle entries with the solution
//      vectors of the group published in LDS: no cross-lane scatter.
// A system whose FP32 factorization is not trusted, because its smallest
// pivot falls below fb_ratio * N * (matrix scale) (condition beyond the
// correction's reach, the Gaussian-family or near-duplicate systems) or its
// drift Schur complement is tiny, is handed to the FP64 kernel, which also
// takes every decision about singular systems.
#pragma once

#include "gs_solve_impl.hpp"

namespace gs {

template <int G, int SRC>
__device__ __forceinline__ float bcastf(float v) {
    return __int_as_float(bcast32<G, SRC>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int PAT>
__device__ __forceinline__ float swz_f(float v) {
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), PAT));
}

// Butterfly sum in FP32; every lane of the group ends with the same value.
template <int G>
__device__ __forceinline__ float gsumf(float v) {
    static_assert(G <= 32, "group size");
    if constexpr (G >= 2) v += dpp_f<0xB1>(v);
    if constexpr (G >= 4) v += dpp_f<0x4E>(v);
    if constexpr (G >= 8) v += swz_f<0x1F | (4 << 10)>(v);
    if constexpr (G >= 16) v += swz_f<0x1F | (8 << 10)>(v);
    if constexpr (G >= 32) v += swz_f<0x1F | (16 << 10)>(v);
    return v;
}

// One lean structure in FP32 (see LeanRegs): the isotropic form keeps
// 1/range and scales each separation component before squaring, so the
// range never meets FP32's exponent limits as range^2.
template <int D>
struct LeanRegs32 {
    float t[D * D];
    float ir, sill, alpha, beta, gam;
    bool aniso, sph, gau;
};

template <int D>
__device__ __forceinline__ LeanRegs32<D> load_lean32(const DevStructure &S) {
    const LeanRegs<D> L = load_lean<D>(S);
    LeanRegs32<D> F;
#pragma unroll
    for (int i = 0; i < D * D; i++) F.t[i] = (float)L.t[i];
    F.ir = (float)S.inv_range;
    F.sill = (float)L.sill;
    F.alpha = (float)L.alpha;
    F.beta = (float)L.beta;
    F.gam = (float)L.gam;
    F.aniso = L.aniso;
    F.sph = L.sph;
    F.gau = L.gau;
    return F;
}

// sill * (1 - g(|T dh| / range)) in FP32 (lean_term in FP64).
template <int D, bool ANISO, bool SPH>
__device__ __forceinline__ float lean_term32(const LeanRegs32<D> &L, const float *dh) {
    float h2s = 0.f;
    if constexpr (ANISO) {
#pragma unroll
        for (int k = 0; k < D; k++) {
            float e = 0.f;
#pragma unroll
            for (int d = 0; d < D; d++) e = fmaf(L.t[k * D + d], dh[d], e);
            h2s = fmaf(e, e, h2s);
        }
    } else {
#pragma unroll
        for (int d = 0; d < D; d++) {
            const float e = dh[d] * L.ir;
            h2s = fmaf(e, e, h2s);
        }
    }
    const float d = __builtin_sqrtf(h2s);
    float c;
    if constexpr (SPH) {
        c = d >= 1.f ? 0.f : 1.f - (1.5f * d - 0.5f * d * d * d);
    } else {
        const float q = L.alpha * d;
        const float p = fmaf(fmaf(L.gam, q, L.beta), q, 1.f);
        c = p * __expf(-(L.gau ? h2s : q));
    }
    return L.sill * c;
}

// FP32 data-data entries e0 <= e < e1 of one pair of row blocks into the
// lane's staging column (same entry numbering as lean_slots); with TGT the
// target entries of the two rows are accumulated into tA / tB.
template <int D, int NB, bool TGT>
__device__ __forceinline__ void staged_entries32(const DevModel &Mdl, const int e0, const int e1,
                                                 const int iA, const double (&xA)[D],
                                                 const double (&xB)[D], const double *xy,
                                                 const double *tg, float *stage, const int lane,
                                                 float &tA, float &tB) {
    float dA[D], dB[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        dA[d] = (float)(xA[d] - tg[d]);
        dB[d] = (float)(xB[d] - tg[d]);
    }
    const int ns = Mdl.ns;
    for (int s = 0; s < ns; s++) {
        const LeanRegs32<D> L = load_lean32<D>(Mdl.s[s]);
        const bool first = s == 0;
#define GS_F32_CASE(AN, SP)                                                                    \
    do {                                                                                       \
        _Pragma("unroll 1") for (int e = e0; e < e1; e++) {                                    \
            const bool inA = e < iA;                                                           \
            const int c = inA ? e : NB - 2 - e;                                                \
            float dh[D];                                                                       \
            _Pragma("unroll") for (int d = 0; d < D; d++)                                     \
                dh[d] = (float)((inA ? xA[d] : xB[d]) - xy[c * D + d]);                        \
            const float v = lean_term32<D, AN, SP>(L, dh);                                     \
            float *slot = stage + (e - e0) * kWave + lane;                                     \
            *slot = first ? v : *slot + v;                                                     \
        }                                                                                      \
        if constexpr (TGT) {                                                                   \
            tA += lean_term32<D, AN, SP>(L, dA);                                               \
            tB += lean_term32<D, AN, SP>(L, dB);                                               \
        }                                                                                      \
    } while (0)
        if (L.aniso) {
            if (L.sph) GS_F32_CASE(true, true);
            else GS_F32_CASE(true, false);
        } else {
            if (L.sph) GS_F32_CASE(false, true);
            else GS_F32_CASE(false, false);
        }
#undef GS_F32_CASE
    }
}

// factor_forward in FP32.  Returns false when a pivot falls below thresh
// (the fallback threshold, far above the singular one).
template <int NB, int G, int Q>
__device__ __forceinline__ bool factor_forward32(float (&a)[NB / G][NB], float (&y)[Q][NB / G],
                                                 const float c0, const float thresh, float *col,
                                                 const int g, float (&invd)[NB / G]) {
    constexpr int R = NB / G;
    int rid[R];
    float dg[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        rid[r] = row_of<G>(r, g);
        dg[r] = c0;
        invd[r] = 0.f;
    }
    bool ok = true;
    static_for<0, NB>([&](auto JJ) {
        constexpr int j = decltype(JJ)::value;
        constexpr int oj = owner_of<G>(j);
        constexpr int rj = j / G;
        const float d = bcastf<G, oj>(dg[rj]);
        ok = ok && (d >= thresh);
        const float inv = __builtin_amdgcn_rsqf(d);
        invd[rj] = (g == oj) ? inv : invd[rj];
        float l[R];
#pragma unroll
        for (int r = rj; r < R; r++) l[r] = a[r][j] * inv;
        l[rj] = (rid[rj] > j) ? l[rj] : 0.f;
#pragma unroll
        for (int r = rj; r < R; r++) dg[r] = fmaf(-l[r], l[r], dg[r]);
#pragma unroll
        for (int r = rj; r < R; r++) col[rid[r]] = l[r];
#pragma unroll
        for (int qq = 0; qq < Q; qq++) col[NB + qq * G + g] = y[qq][rj] * inv;
        wave_sync();
#pragma unroll
        for (int qq = 0; qq < Q; qq++) {
            const float yj = col[NB + qq * G + oj];
#pragma unroll
            for (int r = rj; r < R; r++) y[qq][r] = fmaf(-l[r], yj, y[qq][r]);
        }
        static_for<j + 1, NB - 1>([&](auto KK) {
            constexpr int k = decltype(KK)::value;
            const float lk = col[k];
#pragma unroll
            for (int r = k / G; r < R; r++) a[r][k] = fmaf(-l[r], lk, a[r][k]);
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

// L' x = y for Q right-hand sides (a and invd as left by factor_forward32:
// a[r][c] = L_ic * L_cc below the diagonal); a lane ends with x of its own
// rows.  Column sums over the group's lanes by butterfly.  The sum for
// column j runs over whole row blocks: the entries on and right of the
// diagonal hold finite leftovers of the build and the factorization (the
// caller only trusts systems whose pivots all passed), and they meet the
// x of rows <= j, which are still zero.
template <int NB, int G, int Q>
__device__ __forceinline__ void back_subst32(const float (&a)[NB / G][NB],
                                             const float (&invd)[NB / G],
                                             const float (&y)[Q][NB / G], float (&x)[Q][NB / G],
                                             const int g) {
    constexpr int R = NB / G;
    int rid[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        rid[r] = row_of<G>(r, g);
#pragma unroll
        for (int q = 0; q < Q; q++) x[q][r] = 0.f;
    }
    static_for<0, NB>([&](auto JJ) {
        constexpr int j = NB - 1 - decltype(JJ)::value;
        constexpr int oj = owner_of<G>(j);
        constexpr int rj = j / G;
        float s[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            s[q] = 0.f;
#pragma unroll
            for (int r = rj; r < R; r++) s[q] = fmaf(a[r][j], x[q][r], s[q]);
            s[q] = gsumf<G>(s[q]);
        }
        const bool own = g == oj;
#pragma unroll
        for (int q = 0; q < Q; q++)
            x[q][rj] = own ? (y[q][rj] - invd[rj] * s[q]) * invd[rj] : x[q][rj];
    });
}

// Result of one FP32-refined system.
struct Refined {
    double value;      // z'w (without a simple-kriging mean)
    double var;        // c0 - rhs'K^-1 rhs, unclamped
    double lag;        // mu[0] (ordinary kriging)
    bool ok;           // false: the FP32 factorization is not trusted
};

// One system of up to NB data rows plus M <= 1 drift terms (constant drift,
// ordinary kriging; M = 0 is simple kriging and the Vecchia conditional).
// zrow: the right-hand values of the owned rows (centred for simple
// kriging); rows >= cnt are dead.  xy: the group's neighbour coordinates in
// LDS; tg: the target; sol: NB float4 of LDS for the solution vectors.
template <int D, int NB, int G, int M>
__device__ __forceinline__ Refined fp32r_system(const DevModel &Mdl, const int cnt,
                                                const double (&rowx)[NB / G][D],
                                                const double (&zrow)[NB / G], const double *xy,
                                                const double *tg, float *stage, float *col,
                                                float4 *sol, const int lane, const int g_in,
                                                const double fb_ratio) {
    static_assert(M == 0 || M == 1, "constant drift only");
    constexpr int R = NB / G;
    constexpr int Q = 2 + M;
    constexpr int NE = NB - 1;
    constexpr int HS = stage_slots<NB>();
    constexpr int NPS = (NE + HS - 1) / HS;
    const int g = opaque(g_in);
    const double c0 = Mdl.c0;
    const float c0f = (float)c0;
    int rid[R];
#pragma unroll
    for (int r = 0; r < R; r++) rid[r] = row_of<G>(r, g);

    // ---- FP32 build
    float a[R][NB];
    float y[Q][R];
    static_for<0, (R + 1) / 2>([&](auto PP) {
        constexpr int rA = decltype(PP)::value;
        constexpr int rB = R - 1 - rA;
        constexpr bool single = rA == rB;
        const int iA = rid[rA];
        const bool liveA = iA < cnt;
        const bool liveB = NB - 1 - iA < cnt;
        float tA = 0.f, tB = 0.f;
        static_for<0, NPS>([&](auto PS) {
            constexpr int e0 = decltype(PS)::value * HS;
            constexpr int e1 = e0 + HS < NE ? e0 + HS : NE;
            staged_entries32<D, NB, decltype(PS)::value == 0>(Mdl, e0, e1, iA, rowx[rA], rowx[rB],
                                                               xy, tg, stage, lane, tA, tB);
            static_for<0, G * (rA + 1)>([&](auto CC) {
                constexpr int c = decltype(CC)::value;
                if constexpr (c >= e0 && c < e1)
                    a[rA][c] = liveA ? stage[(c - e0) * kWave + lane] : 0.f;
            });
            if constexpr (!single) {
                static_for<0, G * (rB + 1)>([&](auto CC) {
                    constexpr int c = decltype(CC)::value;
                    constexpr int e = NB - 2 - c;
                    if constexpr (e >= e0 && e < e1)
                        a[rB][c] = liveB ? stage[(e - e0) * kWave + lane] : 0.f;
                });
            }
        });
        if constexpr (G * (rB + 1) == NB) a[rB][NB - 1] = 0.f;
        bool zA = true, zB = true;
#pragma unroll
        for (int d = 0; d < D; d++) {
            zA = zA && rowx[rA][d] == tg[d];
            zB = zB && rowx[rB][d] == tg[d];
        }
        y[0][rA] = liveA ? (zA ? c0f : tA) : 0.f;
        if constexpr (!single) y[0][rB] = liveB ? (zB ? c0f : tB) : 0.f;
    });
#pragma unroll
    for (int r = 0; r < R; r++) {
        const bool live = rid[r] < cnt;
        y[1][r] = live ? (float)zrow[r] : 0.f;
        if constexpr (M == 1) y[2][r] = live ? 1.f : 0.f;
    }
    const double amax = M == 1 ? fmax(c0, 1.0) : c0;
    const float thresh = (float)(fb_ratio * (double)(cnt + M) * amax);
    float invd[R];
    bool ok = factor_forward32<NB, G, Q>(a, y, c0f, thresh, col, g, invd);

    // ---- drift in the forward space, then the three solution vectors
    float mu0 = 0.f, vmu = 0.f, pmu = 0.f;
    constexpr int QB = M == 1 ? 3 : 2;
    float yh[QB][R];
    if constexpr (M == 1) {
        float sff = 0.f, sfb = 0.f, sfz = 0.f;
#pragma unroll
        for (int r = 0; r < R; r++) {
            sff = fmaf(y[2][r], y[2][r], sff);
            sfb = fmaf(y[2][r], y[0][r], sfb);
            sfz = fmaf(y[2][r], y[1][r], sfz);
        }
        const float S = gsumf<G>(sff);
        const float inv_s = 1.f / S;
        mu0 = (gsumf<G>(sfb) - 1.f) * inv_s;
        vmu = gsumf<G>(sfz) * inv_s;
        pmu = -inv_s;
        // a drift Schur complement near the reference's singular test is
        // decided in FP64
        ok = ok && (S >= (float)(1e-9 * amax));
#pragma unroll
        for (int r = 0; r < R; r++) {
            yh[0][r] = fmaf(-y[2][r], mu0, y[0][r]);
            yh[1][r] = fmaf(-y[2][r], vmu, y[1][r]);
            yh[2][r] = y[2][r] * inv_s;
        }
    } else {
#pragma unroll
        for (int r = 0; r < R; r++) {
            yh[0][r] = y[0][r];
            yh[1][r] = y[1][r];
        }
    }
    float x[QB][R];
    back_subst32<NB, G, QB>(a, invd, yh, x, g);

    // ---- FP64 correction
#pragma unroll
    for (int r = 0; r < R; r++) {
        const bool live = rid[r] < cnt;
        sol[rid[r]] = make_float4(live ? x[0][r] : 0.f, live ? x[1][r] : 0.f,
                                  (QB == 3 && live) ? x[QB - 1][r] : 0.f, 0.f);
    }
    wave_sync();
    double w[R], v[R], p[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const bool live = rid[r] < cnt;
        w[r] = live ? (double)x[0][r] : 0.0;
        v[r] = live ? (double)x[1][r] : 0.0;
        p[r] = (QB == 3 && live) ? (double)x[QB - 1][r] : 0.0;
    }
    double zw = 0.0, sw = 0.0, sv = 0.0, sp = 0.0;
    double bww = 0.0, bvw = 0.0, bpw = 0.0;       // w0'A w0, v'A w0, p'A w0
#pragma unroll
    for (int r = 0; r < R; r++) {
        zw = fma(zrow[r], w[r], zw);
        sw += w[r];
        sv += v[r];
        sp += p[r];
        bww = fma(c0 * w[r], w[r], bww);
        bvw = fma(c0 * v[r], w[r], bvw);
        bpw = fma(c0 * p[r], w[r], bpw);
    }
    double tb[R];                                 // FP64 target covariances of the owned rows
#pragma unroll
    for (int r = 0; r < R; r++) tb[r] = 0.0;
    const int ns = Mdl.ns;
    for (int s = 0; s < ns; s++) {
        const LeanRegs<D> L = load_lean<D>(Mdl.s[s]);
#define GS_CORR_CASE(AN, SP)                                                                   \
    do {                                                                                       \
        _Pragma("unroll") for (int r = 0; r < R; r++) {                                        \
            double dh[D];                                                                      \
            _Pragma("unroll") for (int d = 0; d < D; d++) dh[d] = rowx[r][d] - tg[d];          \
            tb[r] += lean_term<D, AN, SP, false>(L, dh);                                              \
        }                                                                                      \
        static_for<0, (R + 1) / 2>([&](auto PP) {                                              \
            constexpr int rA = decltype(PP)::value;                                            \
            constexpr int rB = R - 1 - rA;                                                     \
            const int iA = rid[rA];                                                            \
            const bool liveA = iA < cnt;                                                       \
            const bool liveB = NB - 1 - iA < cnt;                                              \
            const int e_end = rA == rB ? iA : NE;                                              \
            _Pragma("unroll 1") for (int e = 0; e < e_end; e++) {                              \
                const bool inA = e < iA;                                                       \
                const int c = inA ? e : NB - 2 - e;                                            \
                double dh[D];                                                                  \
                _Pragma("unroll") for (int d = 0; d < D; d++)                                 \
                    dh[d] = (inA ? rowx[rA][d] : rowx[rB][d]) - xy[c * D + d];                 \
                double A = lean_term<D, AN, SP, false>(L, dh);                                        \
                A = (inA ? liveA : liveB) ? A : 0.0;                                           \
                const float4 sc = sol[c];                                                      \
                const double wi = inA ? w[rA] : w[rB];                                         \
                const double vi = inA ? v[rA] : v[rB];                                         \
                const double pi = inA ? p[rA] : p[rB];                                         \
                const double wc = sc.x, vc = sc.y, pc = sc.z;                                  \
                bww = fma(A, 2.0 * wi * wc, bww);                                              \
                bvw = fma(A, fma(vi, wc, vc * wi), bvw);                                       \
                if constexpr (QB == 3) bpw = fma(A, fma(pi, wc, pc * wi), bpw);                \
            }                                                                                  \
        });                                                                                    \
    } while (0)
        if (L.aniso) {
            if (L.sph) GS_CORR_CASE(true, true);
            else GS_CORR_CASE(true, false);
        } else {
            if (L.sph) GS_CORR_CASE(false, true);
            else GS_CORR_CASE(false, false);
        }
#undef GS_CORR_CASE
    }
    double bw = 0.0, bv = 0.0, bp = 0.0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        bool z = true;
#pragma unroll
        for (int d = 0; d < D; d++) z = z && rowx[r][d] == tg[d];
        const double b = rid[r] < cnt ? (z ? c0 : tb[r]) : 0.0;
        bw = fma(b, w[r], bw);
        bv = fma(b, v[r], bv);
        bp = fma(b, p[r], bp);
    }
    zw = gsum<G>(zw);
    bw = gsum<G>(bw);
    bv = gsum<G>(bv);
    bww = gsum<G>(bww);
    bvw = gsum<G>(bvw);
    Refined out;
    if constexpr (M == 1) {
        sw = gsum<G>(sw);
        sv = gsum<G>(sv);
        sp = gsum<G>(sp);
        bp = gsum<G>(bp);
        bpw = gsum<G>(bpw);
        const double m0 = mu0, vm = vmu, pm = pmu;
        out.value = zw + (bv + vm - (bvw + m0 * sv + vm * sw));
        out.var = c0 - (2.0 * (bw + m0) - (bww + 2.0 * m0 * sw));
        out.lag = m0 + (bp + pm - (bpw + m0 * sp + pm * sw));
    } else {
        out.value = zw + (bv - bvw);
        out.var = c0 - (2.0 * bw - bww);
        out.lag = __builtin_nan("");
    }
    // a non-finite result (overflow in the FP32 stages) goes to FP64 as well
    out.ok = ok && isfinite(out.value) && isfinite(out.var) && (M == 0 || isfinite(out.lag));
    wave_sync();   // sol and xy are rewritten by the group's next system
    return out;
}

// Ordinary (M = 1) / simple (M = 0) kriging with lean models, one system per
// group of G lanes, grid-strided over the targets or a target list.
template <int D, int NB, int M>
__global__ __launch_bounds__(kWave) void krige_fp32r_kernel(SolveParams P) {
    constexpr int G = fp32r_group(NB);
    constexpr int R = NB / G;
    constexpr int Q = 2 + M;
    constexpr int GPW = kWave / G;
    __shared__ double s_xy[GPW * NB * D];
    __shared__ float s_stage[stage_slots<NB>() * kWave];
    __shared__ __attribute__((aligned(16))) float s_col[GPW * ColStride<NB, Q, G>::value];
    __shared__ float4 s_sol[GPW * NB];
    const int lane = threadIdx.x;
    const int g = lane % G;
    const int grp = lane / G;
    double *xy = s_xy + grp * NB * D;
    float *colbuf = s_col + grp * ColStride<NB, Q, G>::value;
    float4 *sol = s_sol + grp * NB;
    const DevModel &Mdl = *P.model;
    const int64_t nsys = P.list ? (int64_t)(*P.list_len) : P.T;

    for (int64_t sys = xcd_block(blockIdx.x, gridDim.x) * GPW + grp; sys < nsys;
         sys += (int64_t)gridDim.x * GPW) {
        const int64_t t = P.list ? (int64_t)P.list[sys] : sys;
        const int cnt = P.count[t];
        if (cnt < P.min_nb || cnt < M || cnt > NB) {
            if (g == 0) {
                P.value[t] = __builtin_nan("");
                P.variance[t] = __builtin_nan("");
                if (P.lagrange) P.lagrange[t] = __builtin_nan("");
                if (P.status) P.status[t] = cnt > NB ? 4 : (cnt < P.min_nb ? 5 : 1);
            }
            continue;
        }
        double tg[D];
#pragma unroll
        for (int d = 0; d < D; d++) tg[d] = P.targets[t * D + d];
        const int32_t *nbl = P.nbr + t * (int64_t)P.stride;
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
            zr[r] = i < cnt ? (M == 0 ? P.values[nbl[i]] - P.sk_mean : P.values[nbl[i]]) : 0.0;
        }
        wave_sync();
        double rowx[R][D];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = row_of<G>(r, g);
#pragma unroll
            for (int d = 0; d < D; d++) rowx[r][d] = i < cnt ? xy[i * D + d] : 0.0;
        }
        const Refined res = fp32r_system<D, NB, G, M>(Mdl, cnt, rowx, zr, xy, tg, s_stage, colbuf,
                                                      sol, lane, g, P.fb_ratio);
        if (g == 0) {
            if (res.ok) {
                P.value[t] = (M == 0 ? P.sk_mean : 0.0) + res.value;
                P.variance[t] = fmax(res.var, 0.0);
                if (P.lagrange) P.lagrange[t] = res.lag;
                if (P.status) P.status[t] = 0;
            } else {
                P.fb_list[atomicAdd(P.fb_len, 1)] = (int32_t)t;
            }
        }
    }
}

// Vecchia log-likelihood (vecchia_kernel) with FP32-refined conditionals;
// the log-variance and quadratic terms are accumulated in FP64.  A model
// with any untrusted conditional is flagged in refine_fail for the FP64
// kernel.
template <int D, int NB>
__global__ __launch_bounds__(kWave) void vecchia_fp32r_kernel(VecchiaParams P) {
    constexpr int G = fp32r_group(NB);
    constexpr int R = NB / G;
    constexpr int GPW = kWave / G;
    __shared__ double s_xy[GPW * NB * D];
    __shared__ float s_stage[stage_slots<NB>() * kWave];
    __shared__ double s_sum[GPW * kMaxModelsPerLaunch];
    __shared__ __attribute__((aligned(16))) float s_col[GPW * ColStride<NB, 2, G>::value];
    __shared__ float4 s_sol[GPW * NB];
    const int lane = threadIdx.x;
    const int g = lane % G;
    const int grp = lane / G;
    double *xy = s_xy + grp * NB * D;
    float *colbuf = s_col + grp * ColStride<NB, 2, G>::value;
    float4 *sol = s_sol + grp * NB;
    const int nm = P.n_models;
    for (int p = g; p < nm; p += G) s_sum[grp * kMaxModelsPerLaunch + p] = 0.0;
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
            const double sill = Mdl.c0;
            double mu = 0.0, var = sill;
            bool ok = true;
            if (cnt > 0) {
                const Refined res = fp32r_system<D, NB, G, 0>(Mdl, cnt, rowx, zr, xy, tg, s_stage,
                                                              colbuf, sol, lane, g, P.fb_ratio);
                ok = res.ok;
                mu = res.value;
                var = fmax(res.var, 1e-12);
            }
            const double rr = zi - mu;
            const double ll = -0.5 * (ln2pi + log(var) + rr * rr / var);
            if (g == 0) {
                s_sum[grp * kMaxModelsPerLaunch + p] += ll;
                if (!ok) P.refine_fail[p] = 1;
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

template <int D, int NB, int M>
hipError_t launch_krige_fp32r_inst(const SolveParams &p, int64_t grid_groups, hipStream_t s) {
    constexpr int GPW = kWave / fp32r_group(NB);
    int64_t blocks = (grid_groups + GPW - 1) / GPW;
    if (blocks < 1) blocks = 1;
    if (blocks > (1 << 30)) blocks = 1 << 30;
    hipLaunchKernelGGL((krige_fp32r_kernel<D, NB, M>), dim3((unsigned)blocks), dim3(kWave), 0, s, p);
    return hipGetLastError();
}

template <int D, int NB>
hipError_t launch_vecchia_fp32r_inst(const VecchiaParams &p, int64_t n_blocks, hipStream_t s) {
    hipLaunchKernelGGL((vecchia_fp32r_kernel<D, NB>), dim3((unsigned)n_blocks), dim3(kWave), 0, s, p);
    return hipGetLastError();
}

}  // namespace gs
