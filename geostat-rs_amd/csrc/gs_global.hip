// Global-neighbourhood kriging above 64 points (kriging.rs:396-406, 609-612):
// every target conditions on all n data points, so the (n+m)-square system
// is target-independent.  It is built once on the device, factored once by
// LU with partial pivoting in exactly the reference's elimination order
// (lu_factor, linalg.rs:27-77: first maximum pivot, rows swapped whole,
// multipliers stored below the diagonal, rows with a zero multiplier left
// alone), and every target then only forms its right-hand side and runs the
// two triangular substitutions (Lu::solve, linalg.rs:86-109).
//
// Compiled without floating-point contraction: the factorization's
// a - f*b updates round exactly as the reference's.
#include "gs_internal.hpp"

namespace gs {

namespace {

constexpr int kT = 256;

// k-th drift basis value (kriging.rs:410-439) at location x; for external
// drift `xd` holds the normalized covariates of that location.
template <int D>
__device__ __forceinline__ double basis_k(const GlobalParams &P, const double *x,
                                          const double *xd, int k) {
    if (k == 0) return 1.0;
    if (P.method == 3) return xd[k - 1];
    double xn[D];
#pragma unroll
    for (int d = 0; d < D; d++) xn[d] = (x[d] - P.center[d]) / P.scale;
    if (k <= D) return xn[k - 1];
    int idx = 1 + D;
#pragma unroll
    for (int a = 0; a < D; a++)
#pragma unroll
        for (int b = a; b < D; b++) {
            if (idx == k) return xn[a] * xn[b];
            idx++;
        }
    return 0.0;
}

// build_lhs (kriging.rs:527-567) over all points: c0 + err_i on the
// diagonal, c0 - gamma(x_i - x_j) off it, drift rows/columns, zero block.
template <int D, bool LEAN>
__global__ void global_lhs_kernel(GlobalParams P) {
    const int64_t N = P.N;
    const int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (e >= N * N) return;
    const int64_t i = e / N, j = e % N;
    const int64_t n = P.n;
    double v;
    if (i < n && j < n) {
        if (i == j) {
            v = P.model->c0 + (P.err ? P.err[i] : 0.0);
        } else {
            double dh[D];
#pragma unroll
            for (int d = 0; d < D; d++) dh[d] = P.coords[i * D + d] - P.coords[j * D + d];
            v = cov_offdiag<D, LEAN>(*P.model, dh);
        }
    } else if (i < n) {
        v = basis_k<D>(P, P.coords + i * D, P.xdrift ? P.xdrift + i * P.nv : nullptr, (int)(j - n));
    } else if (j < n) {
        v = basis_k<D>(P, P.coords + j * D, P.xdrift ? P.xdrift + j * P.nv : nullptr, (int)(i - n));
    } else {
        v = 0.0;
    }
    P.lu[e] = v;
}

// max |a| over the matrix (the scale of the singular test), one block.
__global__ void absmax_kernel(const double *a, int64_t count, double *out) {
    __shared__ double red[kT];
    double m = 0.0;
    for (int64_t e = threadIdx.x; e < count; e += kT) m = fmax(m, fabs(a[e]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = fmax(red[0], 2.2250738585072014e-308);
}

// Step k, part 1 (one block): first row p >= k with the largest |a[p][k]|
// (strictly larger replaces), the singular test against 1e-12 * scale, the
// row swap, and the multipliers a[i][k] / a[k][k] stored in place.
__global__ void lu_pivot_kernel(double *a, int64_t N, int64_t k, int32_t *piv, const double *scale,
                                int32_t *flag) {
    __shared__ double sv[kT];
    __shared__ int64_t si[kT];
    if (*flag) return;
    const int tid = threadIdx.x;
    double best = -1.0;
    int64_t bi = N;
    for (int64_t i = k + tid; i < N; i += kT) {
        const double v = fabs(a[i * N + k]);
        if (v > best) {   // rows visited in increasing order per thread
            best = v;
            bi = i;
        }
    }
    sv[tid] = best;
    si[tid] = bi;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if (tid < w) {
            const double v2 = sv[tid + w];
            const int64_t i2 = si[tid + w];
            if (v2 > sv[tid] || (v2 == sv[tid] && i2 < si[tid])) {
                sv[tid] = v2;
                si[tid] = i2;
            }
        }
        __syncthreads();
    }
    const double mx = sv[0];
    const int64_t p = si[0];
    if (mx < 1e-12 * scale[0]) {
        if (tid == 0) *flag = 1;
        return;
    }
    if (tid == 0) piv[k] = (int32_t)p;
    if (p != k) {
        for (int64_t j = tid; j < N; j += kT) {
            const double t = a[k * N + j];
            a[k * N + j] = a[p * N + j];
            a[p * N + j] = t;
        }
    }
    __syncthreads();
    const double pivot = a[k * N + k];
    for (int64_t i = k + 1 + tid; i < N; i += kT) a[i * N + k] = a[i * N + k] / pivot;
}

// Step k, part 2: a[i][j] -= f_i * a[k][j] for i, j > k where f_i != 0.
__global__ void lu_update_kernel(double *a, int64_t N, int64_t k, const int32_t *flag) {
    if (*flag) return;
    const int64_t i = k + 1 + blockIdx.y;
    const int64_t j = k + 1 + (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i >= N || j >= N) return;
    const double f = a[i * N + k];
    if (f != 0.0) {
        const double t = f * a[k * N + j];
        a[i * N + j] = a[i * N + j] - t;
    }
}

// Per target (one lane each): right-hand side b0 (covariances to every
// datum, or their block averages, then the target's drift basis), the
// pivot swaps, forward and back substitution with the shared factors, and
// the reference's outputs (kriging.rs:613-644, 747-770).  w (the solution)
// and b0 live in scratch as [row][target] so the lanes' accesses coalesce;
// rows are processed in blocks of RB held in registers.
template <int D, bool LEAN>
__global__ __launch_bounds__(64) void global_solve_kernel(GlobalParams P) {
    constexpr int RB = 16;
    const int64_t tl = (int64_t)blockIdx.x * 64 + threadIdx.x;   // index within the chunk
    const int64_t t = P.t0 + tl;
    const bool live = tl < P.Tc && t < P.T;
    const int64_t N = P.N, n = P.n, S = P.stride;
    const DevModel &Mdl = *P.model;
    const double c0 = Mdl.c0;
    double *w = P.w;
    double *b0 = P.b0;
    const double *lu = P.lu;
    if (live && *P.singular) {
        P.value[t] = __builtin_nan("");
        P.variance[t] = __builtin_nan("");
        if (P.lagrange) P.lagrange[t] = __builtin_nan("");
        if (P.status) P.status[t] = 3;
    }
    if (*P.singular) return;
    double tg[D];
#pragma unroll
    for (int d = 0; d < D; d++) tg[d] = live ? P.targets[t * D + d] : 0.0;
    const int64_t col = live ? tl : 0;
    // right-hand side
    if (live) {
        for (int64_t i = 0; i < n; i++) {
            double bi;
            if (P.n_off > 0) {
                double acc = 0.0;
                for (int u = 0; u < P.n_off; u++) {
                    double dh[D];
#pragma unroll
                    for (int d = 0; d < D; d++)
                        dh[d] = P.coords[i * D + d] - (tg[d] + P.offsets[u * D + d]);
                    acc += cov_offdiag<D, LEAN>(Mdl, dh);
                }
                bi = acc / (double)P.n_off;
            } else {
                double dh[D];
#pragma unroll
                for (int d = 0; d < D; d++) dh[d] = P.coords[i * D + d] - tg[d];
                bi = cov_offdiag<D, LEAN>(Mdl, dh);
            }
            b0[i * S + col] = bi;
        }
        double xd[3] = {0.0, 0.0, 0.0};
        if (P.method == 3)
            for (int q = 0; q < P.nv; q++) xd[q] = (P.tdrift[t * P.nv + q] - P.xmean[q]) / P.xscale[q];
        for (int64_t k = n; k < N; k++) b0[k * S + col] = basis_k<D>(P, tg, xd, (int)(k - n));
        for (int64_t i = 0; i < N; i++) w[i * S + col] = b0[i * S + col];
        // row swaps recorded by the factorization (linalg.rs:91-93)
        for (int64_t k = 0; k < N; k++) {
            const int64_t p = P.piv[k];
            if (p != k) {
                const double x = w[k * S + col];
                w[k * S + col] = w[p * S + col];
                w[p * S + col] = x;
            }
        }
    }
    // forward substitution, unit L (linalg.rs:95-101): per row the terms are
    // subtracted in increasing column order, as in the reference
    for (int64_t i0 = 0; i0 < N; i0 += RB) {
        const int rb = (int)((N - i0) < RB ? (N - i0) : RB);
        double s[RB];
#pragma unroll
        for (int r = 0; r < RB; r++) s[r] = (live && r < rb) ? w[(i0 + r) * S + col] : 0.0;
        for (int64_t j = 0; j < i0; j++) {
            const double wj = live ? w[j * S + col] : 0.0;
#pragma unroll
            for (int r = 0; r < RB; r++)
                if (r < rb) s[r] = s[r] - lu[(i0 + r) * N + j] * wj;
        }
#pragma unroll
        for (int r = 0; r < RB; r++) {
#pragma unroll
            for (int q = 0; q < RB; q++)
                if (q < r && r < rb) s[r] = s[r] - lu[(i0 + r) * N + i0 + q] * s[q];
        }
        if (live)
#pragma unroll
            for (int r = 0; r < RB; r++)
                if (r < rb) w[(i0 + r) * S + col] = s[r];
    }
    // back substitution with U (linalg.rs:102-108), blocks from the bottom;
    // inside a block rows go upwards, each finishing with its division
    for (int64_t i1 = N; i1 > 0; i1 -= RB) {
        const int64_t i0 = i1 - RB > 0 ? i1 - RB : 0;
        const int rb = (int)(i1 - i0);
        double s[RB];
#pragma unroll
        for (int r = 0; r < RB; r++) s[r] = (live && r < rb) ? w[(i0 + r) * S + col] : 0.0;
        for (int64_t j = i1; j < N; j++) {
            const double wj = live ? w[j * S + col] : 0.0;
#pragma unroll
            for (int r = 0; r < RB; r++)
                if (r < rb) s[r] = s[r] - lu[(i0 + r) * N + j] * wj;
        }
#pragma unroll
        for (int r = RB - 1; r >= 0; r--) {
            if (r < rb) {
#pragma unroll
                for (int q = 1; q < RB; q++)
                    if (q > r && q < rb) s[r] = s[r] - lu[(i0 + r) * N + i0 + q] * s[q];
                s[r] = s[r] / lu[(i0 + r) * N + i0 + r];
            }
        }
        if (live)
#pragma unroll
            for (int r = 0; r < RB; r++)
                if (r < rb) w[(i0 + r) * S + col] = s[r];
    }
    if (!live) return;
    // outputs (kriging.rs:615-644; block: 747-770 with C(B,B))
    const double vbase = P.n_off > 0 ? *P.cbb : c0;
    if (P.method == 0) {
        const double mean = P.sk_mean;
        double v = mean, red = 0.0;
        for (int64_t i = 0; i < n; i++) {
            const double wi = w[i * S + col];
            v += wi * (P.values[i] - mean);
            red += wi * b0[i * S + col];
        }
        P.value[t] = v;
        P.variance[t] = fmax(vbase - red, 0.0);
        if (P.lagrange) P.lagrange[t] = __builtin_nan("");
    } else {
        double v = -0.0, red = -0.0;
        for (int64_t i = 0; i < n; i++) v += w[i * S + col] * P.values[i];
        for (int64_t i = 0; i < N; i++) red += w[i * S + col] * b0[i * S + col];
        P.value[t] = v;
        P.variance[t] = fmax(vbase - red, 0.0);
        if (P.lagrange) P.lagrange[t] = w[n * S + col];
    }
    if (P.status) P.status[t] = 0;
}

}  // namespace

hipError_t launch_global_factor(int D, bool lean, GlobalParams p, double *d_scale,
                                hipStream_t s) {
    const int64_t N = p.N;
    const int64_t cnt = N * N;
    const unsigned blocks = (unsigned)((cnt + kT - 1) / kT);
    if (D == 2) {
        if (lean) hipLaunchKernelGGL((global_lhs_kernel<2, true>), dim3(blocks), dim3(kT), 0, s, p);
        else hipLaunchKernelGGL((global_lhs_kernel<2, false>), dim3(blocks), dim3(kT), 0, s, p);
    } else if (D == 3) {
        if (lean) hipLaunchKernelGGL((global_lhs_kernel<3, true>), dim3(blocks), dim3(kT), 0, s, p);
        else hipLaunchKernelGGL((global_lhs_kernel<3, false>), dim3(blocks), dim3(kT), 0, s, p);
    } else {
        return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(absmax_kernel, dim3(1), dim3(kT), 0, s, p.lu, cnt, d_scale);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    for (int64_t k = 0; k < N; k++) {
        hipLaunchKernelGGL(lu_pivot_kernel, dim3(1), dim3(kT), 0, s, p.lu, N, k, p.piv, d_scale,
                           p.singular);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        const int64_t rest = N - k - 1;
        if (rest > 0) {
            hipLaunchKernelGGL(lu_update_kernel, dim3((unsigned)((rest + kT - 1) / kT), (unsigned)rest),
                               dim3(kT), 0, s, p.lu, N, k, p.singular);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

hipError_t launch_global_solve(int D, bool lean, const GlobalParams &p, hipStream_t s) {
    const unsigned blocks = (unsigned)((p.Tc + 63) / 64);
    if (D == 2) {
        if (lean) hipLaunchKernelGGL((global_solve_kernel<2, true>), dim3(blocks), dim3(64), 0, s, p);
        else hipLaunchKernelGGL((global_solve_kernel<2, false>), dim3(blocks), dim3(64), 0, s, p);
    } else if (D == 3) {
        if (lean) hipLaunchKernelGGL((global_solve_kernel<3, true>), dim3(blocks), dim3(64), 0, s, p);
        else hipLaunchKernelGGL((global_solve_kernel<3, false>), dim3(blocks), dim3(64), 0, s, p);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace gs
