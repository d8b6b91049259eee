// Kriging::predict_inner (kriging.rs:569-644) for 33-64 neighbours on the
// matrix cores (FP64 MFMA), with simple, ordinary or universal kriging.
//
// One system per wavefront.  The system is the bordered matrix of 18 x 18
// tiles (4 x 4) in the layout of v_mfma_f64_4x4x4_4b_f64 (lane (x = l >> 4,
// y = l & 3) of block b = (l >> 2) & 3 holds element [y][x] of a tile):
//     [ K  . ]   K: the neighbours' covariances (64 x 64; empty slots are
//     [ B  0 ]      far-away points, c0 on the diagonal and 0 elsewhere)
// where the two border tile rows B hold, as rows, the drift basis at the
// neighbours (F_0 .. F_{M-1}), their covariances to the target (k) and
// their values (z; minus the mean for simple kriging), zero rows after.
// The four MFMA blocks own the tile rows cyclically: block b holds rows
// 4r + b of K (row group r = 0..3, slots kslot(r, j) for j <= 4r + 3; the
// slots right of a block's diagonal are carried along unused) and border
// row 16 + (b & 1) at the even (b < 2) or odd (b >= 2) tile columns.  One
// instruction thus works on four tiles of one system, one per block.
//
// A blocked right-looking Cholesky over the 16 tile columns, carried
// through the border rows, leaves L^-1 [F k z] in them.  Per column q: the
// diagonal tile (block q & 3) is factored and inverted in registers by
// every lane (from its 16 entries, read back through LDS; the pivots are
// tested as the reference's elimination tests them, linalg.rs:36-57), the
// column's tiles are multiplied by L_qq^-T (one MFMA per row group), the
// column is published to LDS, and the trailing tiles X_ij -= L_iq L_jq^T
// are updated with L_jq read from it (one MFMA per row group and column).
// The next diagonal tile is updated and handed on first, so its round trip
// runs under the other updates.  The border rows then go through LDS to one
// neighbour per lane, and the drift block is solved by the modified
// Gram-Schmidt of the vector kernels (drift_solve over a 64-lane group):
// the same value, variance and multiplier formulas and singular tests.
#pragma once

#include <type_traits>

#include "gs_vecchia_mfma_impl.hpp"

namespace gs {

namespace vm64 {

using namespace vm;

constexpr int kNP = 64;       // neighbour slots
constexpr int kBorder = 40;   // slots 40..47: border tile (16 + (b & 1), 2m + (b >> 1))
constexpr int kSlots = 48;
constexpr int kRowS = 66;     // LDS stride of a 64-entry row (+16 B: rows on different banks)
constexpr int kColRows = 20;  // published column: 16 tiles of K, 4 border blocks

// Slot of tile (4r + b, j) of K in block b: row group r keeps 4r + 4 slots.
__host__ __device__ constexpr int kslot(int r, int j) { return 2 * r * (r + 1) + j; }

// One lean structure's covariance at frame distance d (h2 = d^2), with the
// coefficients folded as in Frame (KIND: 0 spherical, 1 Matern 3/2, 2
// Matern 5/2, 3 exponential, 4 Gaussian).
template <int KIND>
__device__ __forceinline__ double lean_cov(double a0, double a1, double a2, double d, double h2) {
    if constexpr (KIND == 4) return a0 * exp_neg256(h2);
    if constexpr (KIND == 0) return d >= 1.0 ? 0.0 : fma(-d, fma(-a2, h2, a1), a0);
    if constexpr (KIND == 1) return fma(a1, d, a0) * exp_neg256(d);
    if constexpr (KIND == 2) return fma(fma(a2, d, a1), d, a0) * exp_neg256(d);
    return a0 * exp_neg256(d);
}

// Covariance of one lean structure, in its frame.
template <int D, bool ANISO, int KIND>
struct Cov1 {
    Frame<D, ANISO, KIND> F;
    __device__ __forceinline__ void init(const DevModel &m) { F = make_frame<D, ANISO, KIND>(m.s[0]); }
    __device__ __forceinline__ void frame(const double *c, double *q) const { to_frame(F, c, q); }
    __device__ __forceinline__ double cov(double h2) const { return cov_h2(F, h2); }
};

// Two lean structures with the same anisotropy (the sum over structures of
// gamma_dh and covariance_dh, model.rs:636-646, 691-693): distances in the first structure's frame, the
// second's distance that one times the ratio of their scales.  KA and KB
// are not Gaussian (both need the distance itself).
template <int D, bool ANISO, int KA, int KB>
struct Cov2 {
    Frame<D, ANISO, KA> F;
    double ratio, b0, b1, b2;
    __device__ __forceinline__ void init(const DevModel &m) {
        F = make_frame<D, ANISO, KA>(m.s[0]);
        const Frame<D, ANISO, KB> G = make_frame<D, ANISO, KB>(m.s[1]);
        const double fa = KA == 1 ? 1.7320508075688772 : (KA == 2 ? 2.23606797749979 : 1.0);
        const double fb = KB == 1 ? 1.7320508075688772 : (KB == 2 ? 2.23606797749979 : 1.0);
        ratio = (fb * m.s[1].inv_range) / (fa * m.s[0].inv_range);
        b0 = G.a0;
        b1 = G.a1;
        b2 = G.a2;
    }
    __device__ __forceinline__ void frame(const double *c, double *q) const { to_frame(F, c, q); }
    __device__ __forceinline__ double cov(double h2) const {
        const double d = sqrt_lean(h2);
        const double db = d * ratio;
        return lean_cov<KA>(F.a0, F.a1, F.a2, d, h2) + lean_cov<KB>(b0, b1, b2, db, db * db);
    }
};

// Covariance class by signature: 0-4 one isotropic lean structure of that
// kind, 8-12 one anisotropic, 16 / 17 exponential + spherical sharing one
// anisotropic / isotropic frame.
template <int D, int SIG>
using cov_of = std::conditional_t<
    (SIG >= 16), Cov2<D, SIG == 16, 3, 0>,
    std::conditional_t<(SIG >= 8), Cov1<D, true, (SIG & 7)>, Cov1<D, false, (SIG & 7)>>>;

}  // namespace vm64

template <int D, class CV, int M>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(2))) void krige_mfma64_kernel(SolveParams P) {
    using namespace vm;
    using namespace vm64;
    static_assert(M >= 0 && M + 2 <= 8, "the border holds M + 2 rows");
    __shared__ double s_q[kNP * D];                  // neighbours in the frame, target at the origin
    __shared__ double s_row[(M + 3) * kRowS];        // rows F_0..F_{M-1}, k, z; a zero row
    __shared__ double s_col[kColRows * 16];          // the published tile column
    __shared__ double s_t[4 * kTS];                  // the next diagonal tile, per block
    __shared__ double s_y[8 * kRowS];                // L^-1 of the border rows
    const int lane = threadIdx.x;
    const int blk = (lane >> 2) & 3;
    const int x_ = lane >> 4, y_ = lane & 3;
    const DevModel &Mdl = *P.model;
    const double c0 = Mdl.c0;
    CV cv;
    cv.init(Mdl);
    // scale of the singular test (linalg.rs:36-39): the bordered matrix's
    // largest entry, max(c0, 1) with a drift (the basis stays in [-1, 1])
    const double thresh = 1e-12 * (M >= 1 ? fmax(c0, 1.0) : c0);
    s_row[(M + 2) * kRowS + lane] = 0.0;
    // the two entries (pi, pj), pi > pj, of a 16 x 16 diagonal super-block
    // this lane evaluates: pairs lane and lane + 64 of the 120 in row-major
    // order (the second clamped to the last pair; its stores are skipped)
    int pi[2], pj[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int pp = min(lane + 64 * h, 119);
        int i = 1;
        while ((i + 1) * i / 2 <= pp) i++;
        pi[h] = i;
        pj[h] = pp - i * (i - 1) / 2;
    }
    load_exp256();
    const int64_t nsys = P.list ? (int64_t)(*P.list_len) : P.T;
    for (int64_t sys = xcd_block(blockIdx.x, gridDim.x); sys < nsys; sys += gridDim.x) {
        const int64_t t = P.list ? (int64_t)P.list[sys] : sys;
        const int craw = P.count[t];
        const bool bad = craw < P.min_nb || craw < M || craw > kNP;
        const int cnt = bad ? 0 : craw;
        double tg[D];
#pragma unroll
        for (int d = 0; d < D; d++) tg[d] = P.targets[t * D + d];
        // one neighbour per lane: its frame position relative to the target
        // (a neighbour on the target at the origin, any other off it; empty
        // slots far out), drift basis, covariance to the target and value
        {
            const int j = lane;
            const bool live = j < cnt;
            const int64_t id = live ? P.nbr[t * (int64_t)P.stride + j] : 0;
            double xa[D], c[D], q[D];
            bool on = true, qz = true;
#pragma unroll
            for (int d = 0; d < D; d++) {
                xa[d] = live ? P.coords[id * D + d] : 0.0;
                c[d] = live ? xa[d] - tg[d] : 0.0;
                on = on && c[d] == 0.0;
            }
            cv.frame(c, q);
#pragma unroll
            for (int d = 0; d < D; d++) qz = qz && q[d] == 0.0;
            if (qz && !on) q[0] = 0x1p-1074;
            if (!live) {
#pragma unroll
                for (int d = 0; d < D; d++) q[d] = d == 0 ? kFar * (j + 1) : 0.0;
            }
            double h2 = 0.0;
#pragma unroll
            for (int d = 0; d < D; d++) {
                s_q[j * D + d] = q[d];
                h2 = fma(q[d], q[d], h2);
            }
            if constexpr (M > 0) {
                double f[M];
                basis<D, M>(P, xa, f);
#pragma unroll
                for (int e = 0; e < M; e++) s_row[e * kRowS + j] = live ? f[e] : 0.0;
            }
            const double kv = cv.cov(h2);
            s_row[M * kRowS + j] = on && live ? c0 : kv;
            const double zv = live ? P.values[id] : 0.0;
            s_row[(M + 1) * kRowS + j] = live ? (M == 0 ? zv - P.sk_mean : zv) : 0.0;
        }
        wave_sync();
        const int y = opaque(y_), x = opaque(x_);
        double X[kSlots];
        // K: tile (4r + blk, j), element [y][x] = C(point 16r + 4 blk + y,
        // point 4j + x).  Left of the diagonal super-block (j < 4r) each lane
        // evaluates its entries; the super-block's 120 distinct entries are
        // evaluated packed over the lanes into a 16 x 17 LDS copy (over the
        // column buffer, idle until the factorization), c0 on its diagonal.
        static_for<0, 4>([&](auto RR) {
            constexpr int r = decltype(RR)::value;
            double *sd = s_col;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int a = 16 * r + pi[h], b = 16 * r + pj[h];
                double h2 = 0.0;
#pragma unroll
                for (int d = 0; d < D; d++) {
                    const double e = s_q[a * D + d] - s_q[b * D + d];
                    h2 = fma(e, e, h2);
                }
                const double v = cv.cov(h2);
                if (h == 0 || lane + 64 < 120) {
                    sd[pi[h] * 17 + pj[h]] = v;
                    sd[pj[h] * 17 + pi[h]] = v;
                }
            }
            if (lane < 16) sd[lane * 18] = c0;
            const int a = 16 * r + 4 * blk + y;
            double pa[D];
#pragma unroll
            for (int d = 0; d < D; d++) pa[d] = s_q[a * D + d];
            static_for<0, 4 * r>([&](auto JJ) {
                constexpr int j = decltype(JJ)::value;
                const int b = 4 * j + x;
                double h2 = 0.0;
#pragma unroll
                for (int d = 0; d < D; d++) {
                    const double e = pa[d] - s_q[b * D + d];
                    h2 = fma(e, e, h2);
                }
                X[kslot(r, j)] = pin(cv.cov(h2));
                if constexpr ((j % 2) == 1) __builtin_amdgcn_sched_barrier(0);
            });
            wave_sync();
            static_for<0, 4>([&](auto CC) {
                constexpr int c = decltype(CC)::value;
                X[kslot(r, 4 * r + c)] = sd[(4 * blk + y) * 17 + 4 * c + x];
            });
            wave_sync();   // the next row group rewrites sd
        });
        // border tile row 16 + rb, tile columns 2m + hb
        const int rb = blk & 1, hb = blk >> 1;
        {
            const int br = 4 * rb + y;
            const double *rrow = br < M + 2 ? s_row + br * kRowS : s_row + (M + 2) * kRowS;
            static_for<0, 8>([&](auto MM) {
                constexpr int m = decltype(MM)::value;
                X[kBorder + m] = rrow[4 * (2 * m + hb) + x];
            });
        }
        bool ok = true;
        s_t[blk * kTS + 4 * y + x] = X[kslot(0, 0)];
        wave_sync();
        double linv = tile_linv<true>(s_t, y, x, thresh, ok);
        static_for<0, 16>([&](auto QQ) {
            constexpr int q = decltype(QQ)::value;
            constexpr int rq = q >> 2;
            // the column below the diagonal: L_iq = X_iq L_qq^-T
            static_for<rq, 4>([&](auto RR) {
                constexpr int r = decltype(RR)::value;
                X[kslot(r, q)] = mfma(linv, X[kslot(r, q)], 0.0);
            });
            {
                // the border tile of column q, in the blocks holding that column
                const double sc = mfma(linv, X[kBorder + (q >> 1)], 0.0);
                X[kBorder + (q >> 1)] = hb == (q & 1) ? pin(sc) : X[kBorder + (q >> 1)];
            }
            if constexpr (q < 15) {
                constexpr int q1 = q + 1, r1 = q1 >> 2;
                static_for<rq, 4>([&](auto RR) {
                    constexpr int r = decltype(RR)::value;
                    s_col[(4 * r + blk) * 16 + 4 * y + x] = X[kslot(r, q)];
                });
                // rows 16 and 17 from the blocks holding column q (the
                // others write rows 18 and 19, never read)
                s_col[(16 + rb + (hb == (q & 1) ? 0 : 2)) * 16 + 4 * y + x] = X[kBorder + (q >> 1)];
                wave_sync();
                // the next diagonal tile first, handed on through LDS
                const double l1 = s_col[q1 * 16 + 4 * y + x];
                X[kslot(r1, q1)] = mfma_nega(l1, X[kslot(r1, q)], X[kslot(r1, q1)]);
                s_t[blk * kTS + 4 * y + x] = X[kslot(r1, q1)];
                // trailing tiles: X_ij -= L_iq L_jq^T (row groups holding rows >= j)
                static_for<q1, 16>([&](auto JJ) {
                    constexpr int j = decltype(JJ)::value;
                    const double l = j == q1 ? l1 : s_col[j * 16 + 4 * y + x];
                    static_for<(j >> 2), 4>([&](auto RR) {
                        constexpr int r = decltype(RR)::value;
                        if constexpr (!(j == q1 && r == r1))
                            X[kslot(r, j)] = mfma_nega(l, X[kslot(r, q)], X[kslot(r, j)]);
                    });
                });
                // border: tile (16 + rb, 2m + hb) -= L_{16+rb,q} L_{2m+hb,q}^T for
                // the columns right of q (at 2m = q only the odd-column blocks)
                const double lb = s_col[(16 + rb) * 16 + 4 * y + x];
                static_for<(q >> 1), 8>([&](auto MM) {
                    constexpr int m = decltype(MM)::value;
                    if constexpr (2 * m + 1 > q) {
                        const double la = s_col[(2 * m + hb) * 16 + 4 * y + x];
                        const double nv = mfma_nega(la, lb, X[kBorder + m]);
                        if constexpr (2 * m > q) X[kBorder + m] = nv;
                        else X[kBorder + m] = hb == 1 ? pin(nv) : X[kBorder + m];
                    }
                });
                wave_sync();
                linv = tile_linv<true>(s_t + (q1 & 3) * kTS, y, x, thresh, ok);
            }
        });
        // L^-1 of the border rows, one neighbour per lane
        if (4 * rb + y < M + 2) {
            static_for<0, 8>([&](auto MM) {
                constexpr int m = decltype(MM)::value;
                s_y[(4 * rb + y) * kRowS + 4 * (2 * m + hb) + x] = X[kBorder + m];
            });
        }
        wave_sync();
        double yv[2 + M][1];
        yv[0][0] = s_y[M * kRowS + lane];
        yv[1][0] = s_y[(M + 1) * kRowS + lane];
#pragma unroll
        for (int e = 0; e < M; e++) yv[2 + e][0] = s_y[e * kRowS + lane];
        double value, var, lag;
        if constexpr (M == 0) {
            value = P.sk_mean + gsum<kWave>(yv[0][0] * yv[1][0]);
            var = c0 - gsum<kWave>(yv[0][0] * yv[0][0]);
            lag = __builtin_nan("");
        } else {
            double fT[M];
            basis<D, M>(P, tg, fT);
            value = drift_solve<1, kWave, M>(yv, fT, c0, thresh, ok, var, lag);
        }
        if (lane == 0) {
            const bool good = !bad && ok;
            P.value[t] = good ? value : __builtin_nan("");
            P.variance[t] = good ? fmax(var, 0.0) : __builtin_nan("");
            if (P.lagrange) P.lagrange[t] = good ? lag : __builtin_nan("");
            if (P.status) P.status[t] = bad ? (craw > kNP ? 4 : (craw < P.min_nb ? 5 : 1)) : (ok ? 0 : 3);
        }
        wave_sync();
    }
}

template <int D, int SIG, int M>
hipError_t launch_krige_mfma64_inst(const SolveParams &p, int64_t grid_groups, hipStream_t s) {
    int64_t blocks = grid_groups;
    if (blocks < 1) blocks = 1;
    if (blocks > (1 << 30)) blocks = 1 << 30;
    hipLaunchKernelGGL((krige_mfma64_kernel<D, vm64::cov_of<D, SIG>, M>), dim3((unsigned)blocks),
                       dim3(kWave), 0, s, p);
    return hipGetLastError();
}

}  // namespace gs
