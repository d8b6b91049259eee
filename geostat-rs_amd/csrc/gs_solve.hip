// Dispatch for the templated solve kernels plus the small helper kernels
// (size bucketing, deterministic Vecchia reduction).  The template bodies
// are in gs_solve_impl.hpp; each (D, NB, M) instantiation is compiled in
// its own generated translation unit (see build_lib.py).

#include "gs_internal.hpp"

namespace gs {

constexpr int kW = 64;

// Size buckets for uneven neighbourhoods (octant search): targets, taken in
// the search's processing order (perm, or index order), are appended to the
// list of the smallest padded size that holds them.  Each wavefront claims
// one run per bucket with a single atomic and keeps its targets' order, so
// a bucket list is a concatenation of 64-target runs in processing order.
__global__ void bucket_kernel(const int32_t *count, const int32_t *perm, int64_t T,
                              const int32_t *n_dev, const uint8_t *skip, int32_t *lists,
                              int32_t *lens, int64_t cap) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = n_dev ? (int64_t)*n_dev : T;
    bool live = i < n;
    const int32_t t = live ? (perm ? perm[i] : (int32_t)i) : 0;
    if (live && skip && skip[t]) live = false;
    const int c = live ? count[t] : 0;
    const int b = c <= 16 ? 0 : (c <= 32 ? 1 : 2);
    const unsigned lane = threadIdx.x & 63u;
    const uint64_t below = lane ? (~0ull >> (64u - lane)) : 0ull;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const uint64_t mask = __ballot(live && b == q);
        if (!mask) continue;
        const int leader = __ffsll((unsigned long long)mask) - 1;
        int base = 0;
        if ((int)lane == leader) base = atomicAdd(&lens[q], __popcll(mask));
        base = __shfl(base, leader, 64);
        if (live && b == q)
            lists[(int64_t)q * cap + base + __popcll(mask & below)] = t;
    }
}

// Per-model sum of the per-block partials in a fixed order: lane-strided
// partial sums, then a fixed pairwise tree.
__global__ void vecchia_reduce_kernel(const double *partial, int64_t n_blocks, int32_t n_models,
                                      double *out) {
    const int p = blockIdx.x;
    const int lane = threadIdx.x;
    double s = 0.0;
    for (int64_t b = lane; b < n_blocks; b += kW) s += partial[b * n_models + p];
    __shared__ double red[kW];
    red[lane] = s;
    __syncthreads();
    for (int w = kW / 2; w > 0; w >>= 1) {
        if (lane < w) red[lane] += red[lane + w];
        __syncthreads();
    }
    if (lane == 0) out[p] = red[0];
}

// Global neighbourhood (kriging.rs:441-444): every target conditions on all
// data points in index order.
__global__ void iota_nbr_kernel(int32_t *nbr, int32_t *count, int64_t T, int32_t n,
                                int32_t stride) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    for (int j = 0; j < stride; j++) nbr[t * stride + j] = j < n ? j : -1;
    count[t] = n;
}

__global__ void fill_i32_kernel(int32_t *p, int64_t n, int32_t value) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

hipError_t launch_iota_nbr(int32_t *nbr, int32_t *count, int64_t T, int32_t n, int32_t stride,
                           hipStream_t s) {
    if (T == 0) return hipSuccess;
    hipLaunchKernelGGL(iota_nbr_kernel, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, s, nbr,
                       count, T, n, stride);
    return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t *p, int64_t n, int32_t value, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n,
                       value);
    return hipGetLastError();
}

hipError_t launch_zero_i32(int32_t *p, int64_t n, hipStream_t s) { return launch_fill_i32(p, n, 0, s); }

__global__ void fill_f64_kernel(double *p, int64_t n, double value) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

hipError_t launch_fill_f64(double *p, int64_t n, double value, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_f64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n,
                       value);
    return hipGetLastError();
}

__global__ void min_i32_kernel(const int32_t *src, int32_t cap, int32_t *dst) {
    if (threadIdx.x == 0) dst[0] = min(src[0], cap);
}

hipError_t launch_min_i32(const int32_t *src, int32_t cap, int32_t *dst, hipStream_t s) {
    hipLaunchKernelGGL(min_i32_kernel, dim3(1), dim3(64), 0, s, src, cap, dst);
    return hipGetLastError();
}

template <int D, int M, int V, bool LEAN, bool X>
static hipError_t krige_nb(int NB, const SolveParams &p, int64_t gg, hipStream_t s) {
    switch (NB) {
    case 16: return launch_krige_inst<D, 16, M, group_size(16, V), LEAN, X>(p, gg, s);
    case 32: return launch_krige_inst<D, 32, M, group_size(32, V), LEAN, X>(p, gg, s);
    case 64: return launch_krige_inst<D, 64, M, group_size(64, V), LEAN, X>(p, gg, s);
    default: return hipErrorInvalidValue;
    }
}

// Extended systems (X) are compiled for group-size variant 0 only.
template <int D, int M, bool X>
static hipError_t krige_v(int NB, int variant, bool lean, const SolveParams &p, int64_t gg,
                          hipStream_t s) {
    if (!lean || X)
        return variant ? hipErrorInvalidDeviceFunction
                       : (lean ? krige_nb<D, M, 0, true, X>(NB, p, gg, s)
                               : krige_nb<D, M, 0, false, X>(NB, p, gg, s));
    return variant ? krige_nb<D, M, 1, true, X>(NB, p, gg, s) : krige_nb<D, M, 0, true, X>(NB, p, gg, s);
}

hipError_t launch_krige_fp64(int D, int NB, int M, int variant, bool lean, bool ext,
                             const SolveParams &p, int64_t grid_groups, hipStream_t s) {
    if (variant < 0 || variant >= kVariants) return hipErrorInvalidValue;
    if (ext) {
        if (D == 2) {
            switch (M) {
            case 0: return krige_v<2, 0, true>(NB, variant, lean, p, grid_groups, s);
            case 1: return krige_v<2, 1, true>(NB, variant, lean, p, grid_groups, s);
            case 2: return krige_v<2, 2, true>(NB, variant, lean, p, grid_groups, s);
            case 3: return krige_v<2, 3, true>(NB, variant, lean, p, grid_groups, s);
            case 4: return krige_v<2, 4, true>(NB, variant, lean, p, grid_groups, s);
            case 6: return krige_v<2, 6, true>(NB, variant, lean, p, grid_groups, s);
            default: return hipErrorInvalidValue;
            }
        } else if (D == 3) {
            switch (M) {
            case 0: return krige_v<3, 0, true>(NB, variant, lean, p, grid_groups, s);
            case 1: return krige_v<3, 1, true>(NB, variant, lean, p, grid_groups, s);
            case 2: return krige_v<3, 2, true>(NB, variant, lean, p, grid_groups, s);
            case 3: return krige_v<3, 3, true>(NB, variant, lean, p, grid_groups, s);
            case 4: return krige_v<3, 4, true>(NB, variant, lean, p, grid_groups, s);
            case 10: return krige_v<3, 10, true>(NB, variant, lean, p, grid_groups, s);
            default: return hipErrorInvalidValue;
            }
        }
        return hipErrorInvalidValue;
    }
    if (D == 2) {
        switch (M) {
        case 0: return krige_v<2, 0, false>(NB, variant, lean, p, grid_groups, s);
        case 1: return krige_v<2, 1, false>(NB, variant, lean, p, grid_groups, s);
        case 3: return krige_v<2, 3, false>(NB, variant, lean, p, grid_groups, s);
        case 6: return krige_v<2, 6, false>(NB, variant, lean, p, grid_groups, s);
        default: return hipErrorInvalidValue;
        }
    } else if (D == 3) {
        switch (M) {
        case 0: return krige_v<3, 0, false>(NB, variant, lean, p, grid_groups, s);
        case 1: return krige_v<3, 1, false>(NB, variant, lean, p, grid_groups, s);
        case 4: return krige_v<3, 4, false>(NB, variant, lean, p, grid_groups, s);
        case 10: return krige_v<3, 10, false>(NB, variant, lean, p, grid_groups, s);
        default: return hipErrorInvalidValue;
        }
    }
    return hipErrorInvalidValue;
}

// Within-block covariance (kriging.rs:729-743): the mean of C(u - v) over
// all pairs of discretization points, with c0 - nugget on the diagonal
// pairs (the nugget is a measure-zero discontinuity of the block
// integral).  One thread, pairs in the reference's order.
template <int D, bool LEAN>
__global__ void block_cbb_kernel(const DevModel *model, const double *offs, int32_t n_off,
                                 double *out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const DevModel &Mdl = *model;
    const double c0c = Mdl.c0 - Mdl.nugget;
    double cbb = 0.0;
    for (int u = 0; u < n_off; u++)
        for (int v = 0; v < n_off; v++) {
            if (u == v) {
                cbb += c0c;
            } else {
                double dh[D];
#pragma unroll
                for (int d = 0; d < D; d++) dh[d] = offs[u * D + d] - offs[v * D + d];
                cbb += cov_offdiag<D, LEAN>(Mdl, dh);
            }
        }
    out[0] = cbb / ((double)n_off * (double)n_off);
}

hipError_t launch_block_cbb(int D, bool lean, const DevModel *model, const double *offs,
                            int32_t n_off, double *out, hipStream_t s) {
    if (D == 2) {
        if (lean) hipLaunchKernelGGL((block_cbb_kernel<2, true>), dim3(1), dim3(64), 0, s, model, offs, n_off, out);
        else hipLaunchKernelGGL((block_cbb_kernel<2, false>), dim3(1), dim3(64), 0, s, model, offs, n_off, out);
    } else if (D == 3) {
        if (lean) hipLaunchKernelGGL((block_cbb_kernel<3, true>), dim3(1), dim3(64), 0, s, model, offs, n_off, out);
        else hipLaunchKernelGGL((block_cbb_kernel<3, false>), dim3(1), dim3(64), 0, s, model, offs, n_off, out);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int D, int M>
static hipError_t krige32r_nb(int NB, const SolveParams &p, int64_t gg, hipStream_t s) {
    switch (NB) {
    case 16: return launch_krige_fp32r_inst<D, 16, M>(p, gg, s);
    case 32: return launch_krige_fp32r_inst<D, 32, M>(p, gg, s);
    case 64: return launch_krige_fp32r_inst<D, 64, M>(p, gg, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_krige_fp32r(int D, int NB, int M, const SolveParams &p, int64_t grid_groups,
                              hipStream_t s) {
    if (D == 2 && M == 0) return krige32r_nb<2, 0>(NB, p, grid_groups, s);
    if (D == 2 && M == 1) return krige32r_nb<2, 1>(NB, p, grid_groups, s);
    if (D == 3 && M == 0) return krige32r_nb<3, 0>(NB, p, grid_groups, s);
    if (D == 3 && M == 1) return krige32r_nb<3, 1>(NB, p, grid_groups, s);
    return hipErrorInvalidValue;
}

hipError_t launch_vecchia_fp32r(int D, int NB, const VecchiaParams &p, int64_t n_blocks,
                                hipStream_t s) {
    if (D == 2) {
        if (NB == 16) return launch_vecchia_fp32r_inst<2, 16>(p, n_blocks, s);
        if (NB == 32) return launch_vecchia_fp32r_inst<2, 32>(p, n_blocks, s);
        if (NB == 64) return launch_vecchia_fp32r_inst<2, 64>(p, n_blocks, s);
    } else if (D == 3) {
        if (NB == 16) return launch_vecchia_fp32r_inst<3, 16>(p, n_blocks, s);
        if (NB == 32) return launch_vecchia_fp32r_inst<3, 32>(p, n_blocks, s);
        if (NB == 64) return launch_vecchia_fp32r_inst<3, 64>(p, n_blocks, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_bucket(const int32_t *count, const int32_t *perm, int64_t T, int32_t *lists,
                         int32_t *lens, int64_t list_cap, hipStream_t s, const int32_t *n_dev,
                         const uint8_t *skip) {
    if (T == 0) return hipSuccess;
    hipLaunchKernelGGL(bucket_kernel, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, s, count,
                       perm, T, n_dev, skip, lists, lens, list_cap);
    return hipGetLastError();
}

template <int D, int V, bool LEAN>
static hipError_t vecchia_nb(int NB, const VecchiaParams &p, int64_t nb, hipStream_t s) {
    switch (NB) {
    case 16: return launch_vecchia_inst<D, 16, group_size(16, V), LEAN>(p, nb, s);
    case 32: return launch_vecchia_inst<D, 32, group_size(32, V), LEAN>(p, nb, s);
    case 64: return launch_vecchia_inst<D, 64, group_size(64, V), LEAN>(p, nb, s);
    default: return hipErrorInvalidValue;
    }
}

template <int D>
static hipError_t vecchia_v(int NB, int variant, bool lean, const VecchiaParams &p, int64_t nb,
                            hipStream_t s) {
    if (!lean) return variant ? hipErrorInvalidDeviceFunction : vecchia_nb<D, 0, false>(NB, p, nb, s);
    return variant ? vecchia_nb<D, 1, true>(NB, p, nb, s) : vecchia_nb<D, 0, true>(NB, p, nb, s);
}

hipError_t launch_vecchia(int D, int NB, int variant, bool lean, const VecchiaParams &p,
                          int64_t n_blocks, hipStream_t s) {
    if (variant < 0 || variant >= kVariants) return hipErrorInvalidValue;
    if (D == 2) return vecchia_v<2>(NB, variant, lean, p, n_blocks, s);
    if (D == 3) return vecchia_v<3>(NB, variant, lean, p, n_blocks, s);
    return hipErrorInvalidValue;
}

// Group sizes of the gradient kernel: two row blocks per lane.
hipError_t launch_vecchia_grad(int D, int NB, const VecchiaParams &p, int64_t n_blocks,
                               hipStream_t s) {
    if (D == 2 && NB == 16) return launch_vecchia_grad_inst<2, 16, 8>(p, n_blocks, s);
    if (D == 2 && NB == 32) return launch_vecchia_grad_inst<2, 32, 16>(p, n_blocks, s);
    if (D == 3 && NB == 16) return launch_vecchia_grad_inst<3, 16, 8>(p, n_blocks, s);
    if (D == 3 && NB == 32) return launch_vecchia_grad_inst<3, 32, 16>(p, n_blocks, s);
    return hipErrorInvalidValue;
}

hipError_t launch_vecchia_reduce(const double *partial, int64_t n_blocks, int32_t n_models,
                                 double *out, hipStream_t s) {
    hipLaunchKernelGGL(vecchia_reduce_kernel, dim3((unsigned)n_models), dim3(kW), 0, s, partial,
                       n_blocks, n_models, out);
    return hipGetLastError();
}

}  // namespace gs
