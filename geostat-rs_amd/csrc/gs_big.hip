// Moving neighbourhoods above 64 points (kriging.rs:441-490 with
// max_neighbors above 64, radius-only searches, large octant quotas).
//
// The register-resident solve kernels hold at most 64 neighbours.  Larger
// systems are solved here, one 256-lane workgroup per target: the bordered
// (n+m)^2 system of build_lhs (kriging.rs:527-567) is built in the
// workgroup's slot of a device scratch, factored by LU with partial pivoting
// in exactly the reference's order (lu_factor, linalg.rs:27-77: first
// maximum pivot, the singular test against 1e-12 of the largest entry, rows
// swapped whole, a - f*b updates skipped where the multiplier is zero), and
// solved by the reference's substitutions (Lu::solve, linalg.rs:86-109) in
// its summation order.  Compiled without contraction, so every operation
// rounds as the reference's does.  The outputs follow predict_inner
// (kriging.rs:569-644) and predict_block (678-777).
#include "gs_internal.hpp"

namespace gs {

namespace {

constexpr int kB = 256;

// k-th drift basis value at x (kriging.rs:410-439); external drift takes
// the normalized covariates xd of that location.
template <int D>
__device__ __forceinline__ double big_basis(const SolveParams &S, int M, const double *x,
                                            const double *xd, int k) {
    if (k == 0) return 1.0;
    if (S.method == 3) return xd[k - 1];
    double xn[D];
#pragma unroll
    for (int d = 0; d < D; d++) xn[d] = (x[d] - S.center[d]) / S.scale;
    if (k <= D) return xn[k - 1];
    int idx = 1 + D;
#pragma unroll
    for (int a = 0; a < D; a++)
#pragma unroll
        for (int b = a; b < D; b++) {
            if (idx == k) return xn[a] * xn[b];
            idx++;
        }
    (void)M;
    return 0.0;
}

template <int D, bool LEAN>
__global__ __launch_bounds__(kB) void big_solve_kernel(BigParams P) {
    const SolveParams &S = P.S;
    const DevModel &Mdl = *S.model;
    const double c0 = Mdl.c0;
    const int tid = threadIdx.x;
    const int M = P.M;
    double *A = P.scratch + (size_t)blockIdx.x * P.slot;
    __shared__ double sv[kB];
    __shared__ int si[kB];
    __shared__ int s_flag;
    for (int64_t pos = blockIdx.x; pos < P.F; pos += gridDim.x) {
        const int64_t t = P.list[pos];
        const int cnt = P.count[pos];
        const int32_t *nb = P.nbr + pos * (int64_t)P.stride;
        int bad = 0;
        if (cnt < S.min_nb) bad = 5;
        else if (cnt < M) bad = 1;
        else if (cnt > P.cap) bad = 4;
        if (bad) {
            if (tid == 0) {
                S.value[t] = __builtin_nan("");
                S.variance[t] = __builtin_nan("");
                if (S.lagrange) S.lagrange[t] = __builtin_nan("");
                if (S.status) S.status[t] = bad;
            }
            continue;
        }
        const int n = cnt;
        const int N = n + M;
        double *b0 = A + (size_t)N * N;      // right-hand side as built
        double *w = b0 + N;                  // solution
        double tg[D];
#pragma unroll
        for (int d = 0; d < D; d++) tg[d] = S.targets[t * D + d];
        double txd[3] = {0.0, 0.0, 0.0};
        if (S.method == 3)
            for (int q = 0; q < P.nv; q++) txd[q] = (S.tdrift[t * P.nv + q] - S.xmean[q]) / S.xscale[q];
        // build_lhs (kriging.rs:527-567) and the right-hand side (613-620)
        for (int64_t e = tid; e < (int64_t)N * N; e += kB) {
            const int i = (int)(e / N), j = (int)(e % N);
            double v;
            if (i < n && j < n) {
                const int64_t gi = nb[i], gj = nb[j];
                if (i == j) {
                    v = c0 + (S.err ? S.err[gi] : 0.0);
                } else {
                    double dh[D];
#pragma unroll
                    for (int d = 0; d < D; d++) dh[d] = S.coords[gi * D + d] - S.coords[gj * D + d];
                    v = cov_offdiag<D, LEAN>(Mdl, dh);
                }
            } else if (i < n) {
                const int64_t gi = nb[i];
                v = big_basis<D>(S, M, S.coords + gi * D, S.xdrift ? S.xdrift + gi * P.nv : nullptr, j - n);
            } else if (j < n) {
                const int64_t gj = nb[j];
                v = big_basis<D>(S, M, S.coords + gj * D, S.xdrift ? S.xdrift + gj * P.nv : nullptr, i - n);
            } else {
                v = 0.0;
            }
            A[e] = v;
        }
        for (int i = tid; i < N; i += kB) {
            double bi;
            if (i < n) {
                const int64_t gi = nb[i];
                if (S.n_off > 0) {
                    double acc = 0.0;
                    for (int u = 0; u < S.n_off; u++) {
                        double dh[D];
#pragma unroll
                        for (int d = 0; d < D; d++)
                            dh[d] = S.coords[gi * D + d] - (tg[d] + S.offsets[u * D + d]);
                        acc += cov_offdiag<D, LEAN>(Mdl, dh);
                    }
                    bi = acc / (double)S.n_off;
                } else {
                    double dh[D];
#pragma unroll
                    for (int d = 0; d < D; d++) dh[d] = S.coords[gi * D + d] - tg[d];
                    bi = cov_offdiag<D, LEAN>(Mdl, dh);
                }
            } else {
                bi = big_basis<D>(S, M, tg, txd, i - n);
            }
            b0[i] = bi;
            w[i] = bi;
        }
        __syncthreads();
        // scale of the singular test: max |entry| (linalg.rs:36-39)
        double m = 0.0;
        for (int64_t e = tid; e < (int64_t)N * N; e += kB) m = fmax(m, fabs(A[e]));
        sv[tid] = m;
        __syncthreads();
        for (int h = kB / 2; h > 0; h >>= 1) {
            if (tid < h) sv[tid] = fmax(sv[tid], sv[tid + h]);
            __syncthreads();
        }
        const double scale = fmax(sv[0], 2.2250738585072014e-308);
        if (tid == 0) s_flag = 0;
        __syncthreads();
        // lu_factor (linalg.rs:42-77); the row swaps are applied to w as
        // they happen, which is the same permutation Lu::solve applies
        // (linalg.rs:91-93)
        for (int k = 0; k < N; k++) {
            double best = -1.0;
            int bi = N;
            for (int i = k + tid; i < N; i += kB) {
                const double v = fabs(A[(size_t)i * N + k]);
                if (v > best) { best = v; bi = i; }
            }
            sv[tid] = best;
            si[tid] = bi;
            __syncthreads();
            for (int h = kB / 2; h > 0; h >>= 1) {
                if (tid < h) {
                    const double v2 = sv[tid + h];
                    const int i2 = si[tid + h];
                    if (v2 > sv[tid] || (v2 == sv[tid] && i2 < si[tid])) { sv[tid] = v2; si[tid] = i2; }
                }
                __syncthreads();
            }
            const double mx = sv[0];
            const int p = si[0];
            __syncthreads();
            if (mx < 1e-12 * scale) {
                if (tid == 0) s_flag = 1;
                break;
            }
            if (p != k) {
                for (int j = tid; j < N; j += kB) {
                    const double x = A[(size_t)k * N + j];
                    A[(size_t)k * N + j] = A[(size_t)p * N + j];
                    A[(size_t)p * N + j] = x;
                }
                if (tid == 0) {
                    const double x = w[k];
                    w[k] = w[p];
                    w[p] = x;
                }
            }
            __syncthreads();
            const double pivot = A[(size_t)k * N + k];
            for (int i = k + 1 + tid; i < N; i += kB) A[(size_t)i * N + k] = A[(size_t)i * N + k] / pivot;
            __syncthreads();
            const int rest = N - k - 1;
            for (int64_t e = tid; e < (int64_t)rest * rest; e += kB) {
                const int i = k + 1 + (int)(e / rest), j = k + 1 + (int)(e % rest);
                const double f = A[(size_t)i * N + k];
                if (f != 0.0) A[(size_t)i * N + j] = A[(size_t)i * N + j] - f * A[(size_t)k * N + j];
            }
            __syncthreads();
        }
        __syncthreads();
        if (s_flag) {
            if (tid == 0) {
                S.value[t] = __builtin_nan("");
                S.variance[t] = __builtin_nan("");
                if (S.lagrange) S.lagrange[t] = __builtin_nan("");
                if (S.status) S.status[t] = 3;
            }
            __syncthreads();
            continue;
        }
        // substitutions (linalg.rs:95-108) and outputs, by one lane in the
        // reference's summation order
        if (tid == 0) {
            for (int i = 0; i < N; i++) {
                double s = w[i];
                for (int j = 0; j < i; j++) s -= A[(size_t)i * N + j] * w[j];
                w[i] = s;
            }
            for (int i = N - 1; i >= 0; i--) {
                double s = w[i];
                for (int j = i + 1; j < N; j++) s -= A[(size_t)i * N + j] * w[j];
                w[i] = s / A[(size_t)i * N + i];
            }
            const double vbase = S.n_off > 0 ? *S.cbb : c0;
            if (S.method == 0) {
                const double mean = S.sk_mean;
                double v = mean, red = 0.0;
                for (int i = 0; i < n; i++) {
                    v += w[i] * (S.values[nb[i]] - mean);
                    red += w[i] * b0[i];
                }
                S.value[t] = v;
                S.variance[t] = fmax(vbase - red, 0.0);
                if (S.lagrange) S.lagrange[t] = __builtin_nan("");
            } else {
                double v = -0.0, red = -0.0;
                for (int i = 0; i < n; i++) v += w[i] * S.values[nb[i]];
                for (int i = 0; i < N; i++) red += w[i] * b0[i];
                S.value[t] = v;
                S.variance[t] = fmax(vbase - red, 0.0);
                if (S.lagrange) S.lagrange[t] = w[n];
            }
            if (S.status) S.status[t] = 0;
        }
        __syncthreads();
    }
}

__global__ void collect_over_kernel(const int32_t *count, int64_t T, int32_t limit, int32_t *list,
                                    int32_t *len) {
    const int64_t t = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (t < T && count[t] > limit) list[atomicAdd(len, 1)] = (int32_t)t;
}

__global__ void max_i32_kernel(const int32_t *v, int64_t n, int32_t *out) {
    __shared__ int32_t red[kB];
    int32_t m = 0;
    for (int64_t i = threadIdx.x; i < n; i += kB) m = max(m, v[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int h = kB / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

}  // namespace

hipError_t launch_big_solve(int D, bool lean, const BigParams &p, int64_t blocks, hipStream_t s) {
    if (p.F == 0) return hipSuccess;
    const dim3 grid((unsigned)blocks);
    if (D == 2) {
        if (lean) hipLaunchKernelGGL((big_solve_kernel<2, true>), grid, dim3(kB), 0, s, p);
        else hipLaunchKernelGGL((big_solve_kernel<2, false>), grid, dim3(kB), 0, s, p);
    } else if (D == 3) {
        if (lean) hipLaunchKernelGGL((big_solve_kernel<3, true>), grid, dim3(kB), 0, s, p);
        else hipLaunchKernelGGL((big_solve_kernel<3, false>), grid, dim3(kB), 0, s, p);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_collect_over(const int32_t *count, int64_t T, int32_t limit, int32_t *list,
                               int32_t *len, hipStream_t s) {
    if (T == 0) return hipSuccess;
    hipLaunchKernelGGL(collect_over_kernel, dim3((unsigned)((T + kB - 1) / kB)), dim3(kB), 0, s,
                       count, T, limit, list, len);
    return hipGetLastError();
}

hipError_t launch_max_i32(const int32_t *counts, int64_t n, int32_t *out, hipStream_t s) {
    hipLaunchKernelGGL(max_i32_kernel, dim3(1), dim3(kB), 0, s, counts, n, out);
    return hipGetLastError();
}

}  // namespace gs
