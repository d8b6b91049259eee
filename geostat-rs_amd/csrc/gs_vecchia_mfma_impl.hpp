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

__host__ __device__ constexpr int tix(int i, int j) { return i < 8 ? i * (i + 1) / 2 + j : 36 + j; }

// The value, pinned where it is computed: the selects of the build then
// stay selects (evaluated in every lane) instead of branches around the
// evaluation, which would split the code into blocks that all carry the
// tiles.
__device__ __forceinline__ double pin(double v) {
    asm volatile("" : "+v"(v));
    return v;
}

__device__ __forceinline__ double mfma(double a, double b, double c) {
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}

// KIND: 0 spherical, 1 Matern 3/2, 2 Matern 5/2, 3 exponential, 4 Gaussian
template <int D, bool ANISO, int KIND>
__device__ __forceinline__ double cov(const LeanRegs<D> &L, double c0, const double *a, const double *b) {
    double dh[D];
    bool zero = true;
#pragma unroll
    for (int d = 0; d < D; d++) {
        dh[d] = a[d] - b[d];
        zero = zero && dh[d] == 0.0;
    }
    const double v = lean_term<D, ANISO, KIND == 0, true, KIND == 0 ? 0 : KIND>(L, dh);
    return zero ? c0 : v;   // gamma(0) = 0
}

// The same covariance for two distinct points: the handles reject
// duplicate coordinates (kriging.rs:312, vecchia.rs:248), and distinct
// doubles have a nonzero difference, so a separation between two data
// points is never the zero vector.
template <int D, bool ANISO, int KIND>
__device__ __forceinline__ double cov_nz(const LeanRegs<D> &L, const double *a, const double *b) {
    double dh[D];
#pragma unroll
    for (int d = 0; d < D; d++) dh[d] = a[d] - b[d];
    return lean_term<D, ANISO, KIND == 0, true, KIND == 0 ? 0 : KIND>(L, dh);
}

// Element [y][x] of L^-1 for the diagonal tile in tt (row-major 4 x 4,
// written by the block): every lane of the block factors the tile
// redundantly, lane 0 of the block stores the ten entries of the inverse
// factor in lt (whose upper triangle is kept at zero), and each lane reads
// its own element back, so the choice is an LDS address rather than a
// chain of lane-dependent selects.  ok is cleared when a pivot fails the
// singular test (GE: pivot >= thresh, else pivot > thresh).
template <bool GE>
__device__ __forceinline__ double tile_linv(const double *tt, double *lt, int g, int y, int x,
                                            double thresh, bool &ok) {
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
    const double v10 = -l10 * i0 * i1;
    const double v21 = -l21 * i1 * i2;
    const double v20 = -(l20 * i0 + l21 * v10) * i2;
    const double v32 = -l32 * i2 * i3;
    const double v31 = -(l31 * i1 + l32 * v21) * i3;
    const double v30 = -((l30 * i0 + l31 * v10) + l32 * v20) * i3;
    if (g == 0) {
        lt[0] = i0;
        lt[4] = v10;
        lt[5] = i1;
        lt[8] = v20;
        lt[9] = v21;
        lt[10] = i2;
        lt[12] = v30;
        lt[13] = v31;
        lt[14] = v32;
        lt[15] = i3;
    }
    wave_sync();
    const double linv = lt[4 * y + x];
    wave_sync();
    return linv;
}

}  // namespace vm

template <int D, bool ANISO, int KIND>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(3))) void vecchia_mfma_kernel(VecchiaParams P) {
    using namespace vm;
    constexpr int NP = 32;                    // conditioning slots per point
    __shared__ double s_xy[4 * NP * D];
    __shared__ double s_rhs[4 * 4 * NP];      // per block: the last tile row's rows k, z, 0, 0
    __shared__ double s_t[4 * 16];
    __shared__ double s_li[4 * 16];           // per block: L^-1 of the current diagonal tile
    __shared__ double s_dg[4 * 48];           // per block: the diagonal tiles' entries below the diagonal
    __shared__ double s_sum[4 * kMaxModelsPerLaunch];
    const int lane = threadIdx.x;
    const int blk = (lane >> 2) & 3;          // the point of this lane
    const int x_ = lane >> 4, y_ = lane & 3;  // element [y][x] of every tile
    double *xy = s_xy + blk * NP * D;
    double *rhs = s_rhs + blk * 4 * NP;
    double *ks = rhs;                         // row 0: covariances to the point
    double *zs = rhs + NP;                    // row 1: centred values of the conditioning set
    double *tt = s_t + blk * 16;
    double *lt = s_li + blk * 16;
    double *dgs = s_dg + blk * 48;
    const int g = x_ * 4 + y_;                // rank of the lane inside its block
    const int nm = P.n_models;
    for (int p = g; p < nm; p += 16) s_sum[blk * kMaxModelsPerLaunch + p] = 0.0;
    for (int j = g; j < 2 * NP; j += 16) rhs[2 * NP + j] = 0.0;
    lt[g] = 0.0;
    load_exp_table();
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
        // slots past the count are zero: every entry evaluated, used or
        // not, comes from finite data
        for (int j = g; j < NP; j += 16) {
            const int64_t id = j < cnt_ ? nbl[j] : 0;
#pragma unroll
            for (int d = 0; d < D; d++) xy[j * D + d] = j < cnt_ ? P.coords[id * D + d] : 0.0;
            zs[j] = j < cnt_ ? P.zc[id] : 0.0;
        }
        wave_sync();
        for (int p = 0; p < nm; p++) {
            const DevModel &Mdl = P.models[p];
            const double c0 = Mdl.c0;
            const LeanRegs<D> L = load_lean<D>(Mdl.s[0]);
            const double thresh = 1e-12 * c0;
            // lane coordinates made opaque per model: the coordinates and
            // masks of the build are then read inside the model loop instead
            // of being hoisted out of it into long-lived registers
            const int y = opaque(y_), x = opaque(x_), cnt = opaque(cnt_);
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
                const double v = pin(cov_nz<D, ANISO, KIND>(L, xy + a * D, xy + b * D));
                dgs[u] = (a < cnt && b < cnt) ? v : 0.0;
            }
            // (the point is not in its own conditioning set)
#pragma unroll
            for (int e2 = 0; e2 < 2; e2++) {
                const int b = 2 * g + e2;
                const double v = pin(cov_nz<D, ANISO, KIND>(L, tg, xy + b * D));
                ks[b] = b < cnt ? v : 0.0;
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
                    const double v = pin(cov_nz<D, ANISO, KIND>(L, pa, pb));
                    X[tix(i, j)] = (a < cnt && b < cnt) ? v : 0.0;
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
                X[tix(8, j)] = rhs[y * NP + 4 * j + x];
            });
            X[tix(8, 8)] = 0.0;
            wave_sync();
            // ---- blocked Cholesky over the eight tile columns of K
            bool ok = true;
            static_for<0, 8>([&](auto PP) {
                constexpr int q = decltype(PP)::value;
                // the diagonal tile, factored by every lane of the block
                tt[4 * y + x] = X[tix(q, q)];
                wave_sync();
                const double linv = tile_linv<false>(tt, lt, g, y, x, thresh, ok);
                // tiles below: L_iq = X_iq L_qq^-T
                static_for<q + 1, 9>([&](auto II) {
                    constexpr int i = decltype(II)::value;
                    X[tix(i, q)] = mfma(linv, X[tix(i, q)], 0.0);
                });
                // trailing tiles: X_ij -= L_iq L_jq^T (the next diagonal first)
                static_for<q + 1, 9>([&](auto II) {
                    constexpr int i = decltype(II)::value;
                    static_for<q + 1, i + 1>([&](auto JJ) {
                        constexpr int j = decltype(JJ)::value;
                        X[tix(i, j)] = mfma(-X[tix(j, q)], X[tix(i, q)], X[tix(i, j)]);
                    });
                });
            });
            // bottom-right tile: [0][0] = -|L^-1 k|^2 (lane x = 0, y = 0),
            // [0][1] = -(L^-1 k).(L^-1 z) (lane x = 1, y = 0)
            const double t00 = __shfl(X[tix(8, 8)], 4 * blk);
            const double t01 = __shfl(X[tix(8, 8)], 16 + 4 * blk);
            if (cnt_ > 0 && g == 0) {
                const double mu = -t01;
                const double var = fmax(c0 + t00, 1e-12);
                const double rr = zi - mu;
                if (live_pt) {
                    s_sum[blk * kMaxModelsPerLaunch + p] += -0.5 * (ln2pi + log(var) + rr * rr / var);
                    if (!ok) atomicMax(&P.status[p], 3);
                }
            } else if (g == 0 && live_pt) {
                const double rr = zi;
                s_sum[blk * kMaxModelsPerLaunch + p] += -0.5 * (ln2pi + log(c0) + rr * rr / c0);
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
    __shared__ double s_xy[4 * NP * D];
    __shared__ double s_rhs[4 * 4 * NP];      // per block: the last tile row's rows f, k, z, 0
    __shared__ double s_t[4 * 16];
    __shared__ double s_li[4 * 16];
    __shared__ double s_dg[4 * 48];
    const int lane = threadIdx.x;
    const int blk = (lane >> 2) & 3;
    const int x_ = lane >> 4, y_ = lane & 3;
    const int g = x_ * 4 + y_;
    double *xy = s_xy + blk * NP * D;
    double *rhs = s_rhs + blk * 4 * NP;
    double *fs = rhs, *ks = rhs + NP, *vs = rhs + 2 * NP;
    double *tt = s_t + blk * 16;
    double *lt = s_li + blk * 16;
    double *dgs = s_dg + blk * 48;
    const DevModel &Mdl = *P.model;
    const double c0 = Mdl.c0;
    const LeanRegs<D> L = load_lean<D>(Mdl.s[0]);
    const double thresh = 1e-12 * (M >= 1 ? fmax(c0, 1.0) : c0);
    for (int j = g; j < NP; j += 16) rhs[3 * NP + j] = 0.0;
    lt[g] = 0.0;
    load_exp_table();
    wave_sync();
    const int64_t nsys = P.list ? (int64_t)(*P.list_len) : P.T;
    for (int64_t sb = (int64_t)blockIdx.x * 4; sb < nsys; sb += (int64_t)gridDim.x * 4) {
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
        for (int j = g; j < NP; j += 16) {
            const int64_t id = j < cnt_ ? nbl[j] : 0;
#pragma unroll
            for (int d = 0; d < D; d++) xy[j * D + d] = j < cnt_ ? P.coords[id * D + d] : 0.0;
            vs[j] = j < cnt_ ? (M == 0 ? P.values[id] - P.sk_mean : P.values[id]) : 0.0;
            fs[j] = (M == 1 && j < cnt_) ? 1.0 : 0.0;
        }
        wave_sync();
        const int y = opaque(y_), x = opaque(x_), cnt = opaque(cnt_);
        double X[kTiles];
#pragma unroll
        for (int e3 = 0; e3 < 3; e3++) {
            const int u = 3 * g + e3;
            const int tt6 = u / 6, e = u % 6;
            const int r = e >= 3 ? 3 : (e >= 1 ? 2 : 1);
            const int c = e - r * (r - 1) / 2;
            const int a = 4 * tt6 + r, b = 4 * tt6 + c;
            const double v = pin(cov_nz<D, ANISO, KIND>(L, xy + a * D, xy + b * D));
            dgs[u] = (a < cnt && b < cnt) ? v : 0.0;
        }
        // a target may sit on a datum: c0 there
#pragma unroll
        for (int e2 = 0; e2 < 2; e2++) {
            const int b = 2 * g + e2;
            const double v = pin(cov<D, ANISO, KIND>(L, c0, tg, xy + b * D));
            ks[b] = b < cnt ? v : 0.0;
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
                const double v = pin(cov_nz<D, ANISO, KIND>(L, pa, pb));
                X[tix(i, j)] = (a < cnt && b < cnt) ? v : 0.0;
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
            X[tix(8, j)] = rhs[y * NP + 4 * j + x];
        });
        X[tix(8, 8)] = 0.0;
        wave_sync();
        bool ok = true;
        static_for<0, 8>([&](auto PP) {
            constexpr int q = decltype(PP)::value;
            tt[4 * y + x] = X[tix(q, q)];
            wave_sync();
            const double linv = tile_linv<true>(tt, lt, g, y, x, thresh, ok);
            static_for<q + 1, 9>([&](auto II) {
                constexpr int i = decltype(II)::value;
                X[tix(i, q)] = mfma(linv, X[tix(i, q)], 0.0);
            });
            static_for<q + 1, 9>([&](auto II) {
                constexpr int i = decltype(II)::value;
                static_for<q + 1, i + 1>([&](auto JJ) {
                    constexpr int j = decltype(JJ)::value;
                    X[tix(i, j)] = mfma(-X[tix(j, q)], X[tix(i, q)], X[tix(i, j)]);
                });
            });
        });
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
