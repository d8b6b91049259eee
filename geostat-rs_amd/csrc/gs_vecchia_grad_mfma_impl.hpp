// loglik_grad_with_plan (vecchia.rs:416-566) for batches of one isotropic
// structure of a g_and_dg kind, with the factorizations on the matrix cores.
//
// The layout and factorization are vecchia_mfma_kernel's (four ordered
// points per wavefront, 9 x 9 tiles, factor_tiles), with three right-hand
// side rows: r (correlations to the point), z and rho (d g'(d) to the
// point), so the Gram tile holds |L^-1 r|^2, (L^-1 r).(L^-1 z),
// rho.K^-1 r and rho.K^-1 z.  K^-1 r and K^-1 z themselves come from a
// tile back substitution of the first two rows (W_q = (Y_q - sum_{i>q} W_i
// L_iq) L_qq^-1, one MFMA per tile with L_iq transposed through a lane
// shuffle), and a second pass over the lane's pair entries forms the
// bilinear forms with P_ab = d_ab g'(d_ab).  The per-point algebra is
// vecchia_grad_kernel's.
#pragma once

#include "gs_vecchia_mfma_impl.hpp"

namespace gs {

namespace vm {

// Correlation and d g'(d) from the squared distance in range units, as
// grad_pair computes them (KIND as the matrix-core kinds).
template <int KIND>
__device__ __forceinline__ void grad_h2(double h2s, double &c, double &rho) {
    const double d = sqrt_lean(h2s);
    if constexpr (KIND == 0) {
        const bool in = d < 1.0;
        c = in ? 1.0 - (1.5 * d - 0.5 * d * d * d) : 0.0;
        rho = in ? d * (1.5 - 1.5 * d * d) : 0.0;
    } else {
        constexpr double alpha = KIND == 1 ? 1.7320508075688772 : (KIND == 2 ? 2.23606797749979 : 1.0);
        constexpr double beta = (KIND == 1 || KIND == 2) ? 1.0 : 0.0;
        constexpr double gam = KIND == 2 ? 1.0 / 3.0 : 0.0;
        constexpr double r1 = KIND == 3 ? 1.0 : 0.0;
        constexpr double r2 = KIND == 4 ? 2.0 : (KIND == 1 ? 3.0 : (KIND == 2 ? 5.0 / 3.0 : 0.0));
        constexpr double r3 = KIND == 2 ? 5.0 * 2.23606797749979 / 3.0 : 0.0;
        const double q = alpha * d;
        const double e = exp_neg256(KIND == 4 ? h2s : q);
        c = fma(fma(gam, q, beta), q, 1.0) * e;
        rho = d * fma(fma(r3, d, r2), d, r1) * e;
    }
}

// Sum over the 16 lanes of a block (lanes differing in bits 0-1 and 4-5).
__device__ __forceinline__ double block_sum(double v) {
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}

}  // namespace vm

template <int D, int KIND>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(2))) void vecchia_grad_mfma_kernel(VecchiaParams P) {
    using namespace vm;
    constexpr int NP = 32;
    constexpr int kQ = 4;                      // loglik, d/dnugget, d/dpsill, d/drange
    __shared__ double s_xy[4 * kXYS<D>];       // per block: the conditioning set relative to the point
    __shared__ double s_q[4 * kXYS<D>];        // the same over the range
    __shared__ double s_rhs[4 * 3 * NP + NP];  // per block: rows r, z, rho; a zero row
    __shared__ double s_w[4 * 2 * NP];         // per block: K^-1 r and K^-1 z by slot
    __shared__ double s_t[4 * kTS];
    __shared__ double s_dg[4 * kDGS];
    __shared__ double s_sum[4 * kMaxModelsPerLaunch * kQ];
    const int lane = threadIdx.x;
    const int blk = (lane >> 2) & 3;
    const int x_ = lane >> 4, y_ = lane & 3;
    const int g = x_ * 4 + y_;
    double *xy = s_xy + blk * kXYS<D>;
    double *qs = s_q + blk * kXYS<D>;
    double *rhs = s_rhs + blk * 3 * NP;
    double *ks = rhs, *zs = rhs + NP, *rs = rhs + 2 * NP;
    const double *rrow = y_ < 3 ? rhs + y_ * NP : s_rhs + 12 * NP;
    double *ws = s_w + blk * 2 * NP;
    double *tt = s_t + blk * kTS;
    double *dgs = s_dg + blk * kDGS;
    const int nm = P.n_models;
    for (int c = g; c < nm * kQ; c += 16) s_sum[blk * kMaxModelsPerLaunch * kQ + c] = 0.0;
    if (lane < NP) s_rhs[12 * NP + lane] = 0.0;
    load_exp256();
    wave_sync();
    const double ln2pi = 1.8378770664093453;
    // the lane that holds element [x][y] of a tile (its transpose's [y][x])
    const int tlane = 16 * y_ + 4 * blk + x_;

    const int64_t k0 = (int64_t)blockIdx.x * P.points_per_block;
    int64_t k1 = k0 + P.points_per_block;
    if (k1 > P.n) k1 = P.n;
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
        for (int j = g; j < NP; j += 16) {
            const int64_t id = j < cnt_ ? nbl[j] : 0;
#pragma unroll
            for (int d = 0; d < D; d++) xy[j * D + d] = j < cnt_ ? P.coords[id * D + d] - tg[d] : 0.0;
            zs[j] = j < cnt_ ? P.zc[id] : 0.0;
        }
        wave_sync();
        for (int p = 0; p < nm; p++) {
            const DevModel &Mdl = P.models[p];
            const double nug = Mdl.nugget;
            const double sill = Mdl.c0;
            const double psill = Mdl.s[0].sill;
            const double ir = Mdl.s[0].inv_range;
            double ll, gn, gp, gr;
            bool ok = true;
            // the set over the range (empty slots far out)
            for (int j = g; j < NP; j += 16) {
#pragma unroll
                for (int d = 0; d < D; d++)
                    qs[j * D + d] = j < cnt_ ? xy[j * D + d] * ir : (d == 0 ? kFar * (j + 1) : 0.0);
            }
            wave_sync();
            const int y = opaque(y_), x = opaque(x_);
            double X[kTiles];
#pragma unroll
            for (int e3 = 0; e3 < 3; e3++) {
                const int u = 3 * g + e3;
                const int t = u / 6, e = u % 6;
                const int r = e >= 3 ? 3 : (e >= 1 ? 2 : 1);
                const int c = e - r * (r - 1) / 2;
                const int a = 4 * t + r, b = 4 * t + c;
                double h2 = 0.0;
#pragma unroll
                for (int d = 0; d < D; d++) {
                    const double dd = qs[a * D + d] - qs[b * D + d];
                    h2 = fma(dd, dd, h2);
                }
                double cv, rv;
                grad_h2<KIND>(h2, cv, rv);
                dgs[u] = psill * cv;
            }
#pragma unroll
            for (int e2 = 0; e2 < 2; e2++) {
                const int b = 2 * g + e2;
                double h2 = 0.0;
#pragma unroll
                for (int d = 0; d < D; d++) h2 = fma(qs[b * D + d], qs[b * D + d], h2);
                double cv, rv;
                grad_h2<KIND>(h2, cv, rv);
                ks[b] = cv;
                rs[b] = rv;
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
                    double h2 = 0.0;
#pragma unroll
                    for (int d = 0; d < D; d++) {
                        const double dd = pa[d] - qs[b * D + d];
                        h2 = fma(dd, dd, h2);
                    }
                    double cv, rv;
                    grad_h2<KIND>(h2, cv, rv);
                    X[tix(i, j)] = pin(psill * cv);
                    if constexpr ((tix(i, j) % 2) == 1) __builtin_amdgcn_sched_barrier(0);
                });
            });
            wave_sync();
            {
                const int hi = y > x ? y : x, lo = y > x ? x : y;
                const int e = hi * (hi - 1) / 2 + lo;
                static_for<0, 8>([&](auto TT) {
                    constexpr int t = decltype(TT)::value;
                    X[tix(t, t)] = y == x ? sill : dgs[6 * t + e];
                });
            }
            static_for<0, 8>([&](auto JJ) {
                constexpr int j = decltype(JJ)::value;
                X[tix(8, j)] = rrow[4 * j + x];
            });
            X[tix(8, 8)] = 0.0;
            wave_sync();
            factor_tiles<false>(X, tt, y, x, 1e-12 * sill, ok);
            // Gram entries (negated): [0][0] |L^-1 r|^2, [0][1] (L^-1 r).(L^-1 z),
            // [0][2] rho.K^-1 r, [1][2] rho.K^-1 z
            const double yr2 = -__shfl(X[tix(8, 8)], 4 * blk);
            const double yrz = -__shfl(X[tix(8, 8)], 16 + 4 * blk);
            const double rw = -__shfl(X[tix(8, 8)], 32 + 4 * blk);
            const double rvv = -__shfl(X[tix(8, 8)], 32 + 4 * blk + 1);
            // back substitution, rows 0-1 (K^-1 r, K^-1 z as rows by slot)
            double W[8];
            bool okd = true;
            static_for<0, 8>([&](auto QQ) {
                constexpr int q = 7 - decltype(QQ)::value;
                double T = X[tix(8, q)];
                static_for<q + 1, 8>([&](auto II) {
                    constexpr int i = decltype(II)::value;
                    T = mfma_nega(__shfl(X[tix(i, q)], tlane), W[i], T);
                });
                tt[4 * y + x] = X[tix(q, q)];
                wave_sync();
                const double linvT = tile_linv<false>(tt, x, y, 0.0, okd);
                wave_sync();
                W[q] = mfma(linvT, T, 0.0);
            });
            if (y < 2) {
                static_for<0, 8>([&](auto QQ) {
                    constexpr int q = decltype(QQ)::value;
                    ws[y * NP + 4 * q + x] = W[q];
                });
            }
            wave_sync();
            // |K^-1 r|^2 and (K^-1 r).(K^-1 z) over the lane's two slots
            double w2 = 0.0, wvv = 0.0;
            for (int j = g; j < NP; j += 16) {
                const double wa = ws[j], va = ws[NP + j];
                w2 = fma(wa, wa, w2);
                wvv = fma(wa, va, wvv);
            }
            // the bilinear forms over the lane's pair entries (each pair of
            // distinct slots once: 3 diagonal-tile entries, 28 tile entries)
            double b1 = 0.0, b2 = 0.0;
#pragma unroll
            for (int e3 = 0; e3 < 3; e3++) {
                const int u = 3 * g + e3;
                const int t = u / 6, e = u % 6;
                const int r = e >= 3 ? 3 : (e >= 1 ? 2 : 1);
                const int c = e - r * (r - 1) / 2;
                const int a = 4 * t + r, b = 4 * t + c;
                double h2 = 0.0;
#pragma unroll
                for (int d = 0; d < D; d++) {
                    const double dd = qs[a * D + d] - qs[b * D + d];
                    h2 = fma(dd, dd, h2);
                }
                double cv, rv;
                grad_h2<KIND>(h2, cv, rv);
                const double wa = ws[a], wb = ws[b], va = ws[NP + a], vb = ws[NP + b];
                b1 = fma(rv, wa * wb, b1);
                b2 = fma(rv, fma(wa, vb, wb * va), b2);
            }
            static_for<1, 8>([&](auto II) {
                constexpr int i = decltype(II)::value;
                const int a = 4 * i + y;
                double pa[D];
#pragma unroll
                for (int d = 0; d < D; d++) pa[d] = qs[a * D + d];
                const double wa = ws[a], va = ws[NP + a];
                static_for<0, i>([&](auto JJ) {
                    constexpr int j = decltype(JJ)::value;
                    const int b = 4 * j + x;
                    double h2 = 0.0;
#pragma unroll
                    for (int d = 0; d < D; d++) {
                        const double dd = pa[d] - qs[b * D + d];
                        h2 = fma(dd, dd, h2);
                    }
                    double cv, rv;
                    grad_h2<KIND>(h2, cv, rv);
                    const double wb = ws[b], vb = ws[NP + b];
                    b1 = fma(rv, wa * wb, b1);
                    b2 = fma(rv, fma(wa, vb, wb * va), b2);
                });
            });
            w2 = block_sum(w2);
            wvv = block_sum(wvv);
            b1 = 2.0 * block_sum(b1);
            b2 = block_sum(b2);
            wave_sync();   // the set, the rows and the vectors are rewritten next
            if (cnt_ == 0) {
                // first point: N(0, sill); only nugget and psill move sill
                const double var = sill, r = zi;
                ll = -0.5 * (ln2pi + log(var) + r * r / var);
                const double common = -0.5 / var + 0.5 * r * r / (var * var);
                gn = common;
                gp = common;
                gr = 0.0;
            } else {
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
            if (g == 0 && live_pt) {
                double *sums = s_sum + blk * kMaxModelsPerLaunch * kQ;
                sums[p * kQ + 0] += ll;
                sums[p * kQ + 1] += gn;
                sums[p * kQ + 2] += gp;
                sums[p * kQ + 3] += gr;
                if (cnt_ > 0 && !ok) atomicMax(&P.status[p], 3);
            }
        }
        wave_sync();
    }
    __syncthreads();
    for (int c = lane; c < nm * kQ; c += kWave) {
        double acc = 0.0;
        for (int q = 0; q < 4; q++) acc += s_sum[q * kMaxModelsPerLaunch * kQ + c];
        P.partial[(int64_t)blockIdx.x * nm * kQ + c] = acc;
    }
}

template <int D, int KIND>
hipError_t launch_vecchia_grad_mfma_inst(const VecchiaParams &p, int64_t n_blocks, hipStream_t s) {
    hipLaunchKernelGGL((vecchia_grad_mfma_kernel<D, KIND>), dim3((unsigned)n_blocks), dim3(kWave), 0,
                       s, p);
    return hipGetLastError();
}

}  // namespace gs
