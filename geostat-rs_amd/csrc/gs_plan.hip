// Nearest predecessors of the Vecchia plan on the GPU (vecchia.rs:262-281 with
// BucketGrid::k_nearest, search.rs:221-313): for every ordered position k,
// the m nearest points among positions < k.
//
// One lane per query.  The lane walks the same rings of the same bucket grid
// as the reference and stops by the same rule (ring r is skipped once
// ((r-1) cell)^2 exceeds the current m-th distance), so it examines the same
// candidates; it keeps the canonical "m smallest by (d2, position)" in an
// LDS heap.  The reference's heap accepts a candidate into a full heap only
// when its d2 is strictly smaller, so it agrees with the canonical set
// unless a candidate left out has d2 equal to the m-th distance: such
// queries are flagged and the host replays them in the reference's visiting
// order.  Positions below a small threshold scan all predecessors directly
// (the same candidate set: every ring is visited while the heap is not
// full).  Compiled without contraction so d2 rounds as the reference's.
#include "gs_internal.hpp"

namespace gs {

namespace {

constexpr int kL = 64;

struct PredHeap {
    double *d;
    int32_t *p;
    int lane;
    __device__ double &dd(int e) { return d[e * kL + lane]; }
    __device__ int32_t &pp(int e) { return p[e * kL + lane]; }
    __device__ static bool gt(double a, int32_t ai, double b, int32_t bi) {
        return a > b || (a == b && ai > bi);
    }
    __device__ void sift_up(int e) {
        while (e > 0) {
            const int par = (e - 1) >> 1;
            if (!gt(dd(e), pp(e), dd(par), pp(par))) break;
            const double td = dd(e);
            const int32_t tp = pp(e);
            dd(e) = dd(par);
            pp(e) = pp(par);
            dd(par) = td;
            pp(par) = tp;
            e = par;
        }
    }
    __device__ void sift_down(int len, int e) {
        for (;;) {
            const int l = 2 * e + 1, r = l + 1;
            int g = e;
            if (l < len && gt(dd(l), pp(l), dd(g), pp(g))) g = l;
            if (r < len && gt(dd(r), pp(r), dd(g), pp(g))) g = r;
            if (g == e) break;
            const double td = dd(e);
            const int32_t tp = pp(e);
            dd(e) = dd(g);
            pp(e) = pp(g);
            dd(g) = td;
            pp(g) = tp;
            e = g;
        }
    }
};

template <int D>
__device__ __forceinline__ double dist2(const double *a, const double *b) {
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < D; d++) {
        const double v = a[d] - b[d];
        s += v * v;
    }
    return s;
}

__device__ __forceinline__ int64_t sat_i64(double v) {
    if (v != v) return 0;
    if (v >= 9.2233720368547758e18) return INT64_MAX;
    if (v <= -9.2233720368547758e18) return INT64_MIN;
    return (int64_t)v;
}

template <int D>
__global__ __launch_bounds__(kL) void pred_knn_kernel(PredParams P) {
    extern __shared__ unsigned char smem[];
    double *hd = reinterpret_cast<double *>(smem);
    int32_t *hp = reinterpret_cast<int32_t *>(smem + sizeof(double) * P.m * kL);
    const int lane = threadIdx.x;
    const int64_t k = (int64_t)blockIdx.x * kL + lane;
    if (k >= P.n) return;
    PredHeap H{hd, hp, lane};
    const int m = P.m;
    double q[D];
#pragma unroll
    for (int d = 0; d < D; d++) q[d] = P.pts[k * D + d];
    int len = 0;
    double rej = __builtin_inf();   // smallest d2 left out of the heap
    auto push = [&](double d2, int32_t pos) {
        if (len < m) {
            H.dd(len) = d2;
            H.pp(len) = pos;
            H.sift_up(len);
            len++;
        } else if (PredHeap::gt(H.dd(0), H.pp(0), d2, pos)) {
            rej = fmin(rej, H.dd(0));
            H.dd(0) = d2;
            H.pp(0) = pos;
            H.sift_down(len, 0);
        } else {
            rej = fmin(rej, d2);
        }
    };
    if (k <= P.brute) {
        for (int64_t j = 0; j < k; j++) push(dist2<D>(P.pts + j * D, q), (int32_t)j);
    } else {
        int64_t tc[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < D; d++) {
            int64_t v = sat_i64((q[d] - P.origin[d]) / P.cell);
            tc[d] = v < 0 ? 0 : (v > P.nn[d] - 1 ? P.nn[d] - 1 : v);
        }
        int64_t max_ring = P.nn[0];
#pragma unroll
        for (int d = 1; d < D; d++) max_ring = max_ring > P.nn[d] ? max_ring : P.nn[d];
        auto visit = [&](const int64_t *cur) {
            int64_t b = 0;
#pragma unroll
            for (int d = 0; d < D; d++) b += cur[d] * P.strides[d];
            const int64_t s0 = P.start[b], s1 = P.start[b + 1];
            for (int64_t t = s0; t < s1; t++) {
                const int32_t pos = P.items[t];
                if (pos >= k) break;   // positions ascend inside a bucket
                push(dist2<D>(P.pts + (int64_t)pos * D, q), pos);
            }
        };
        for (int64_t ring = 0; ring <= max_ring; ring++) {
            const double rm = fmax((double)ring - 1.0, 0.0) * P.cell;
            const double worst = len == m ? H.dd(0) : __builtin_inf();
            if (rm * rm > worst) break;
            if (ring == 0) {
                visit(tc);
                continue;
            }
            const int64_t r = ring;
            for (int fd = 0; fd < D; fd++) {
                for (int sd = 0; sd < 2; sd++) {
                    const int64_t side = sd == 0 ? -r : r;
                    int64_t lo[3], hi[3];
                    bool empty = false;
#pragma unroll
                    for (int d = 0; d < D; d++) {
                        if (d < fd) { lo[d] = tc[d] - r + 1; hi[d] = tc[d] + r - 1; }
                        else if (d == fd) { lo[d] = tc[d] + side; hi[d] = tc[d] + side; }
                        else { lo[d] = tc[d] - r; hi[d] = tc[d] + r; }
                        // cells outside the grid hold nothing
                        if (lo[d] < 0) lo[d] = 0;
                        if (hi[d] > P.nn[d] - 1) hi[d] = P.nn[d] - 1;
                        if (lo[d] > hi[d]) empty = true;
                    }
                    if (empty) continue;
                    int64_t cur[3] = {lo[0], lo[1], lo[2]};
                    for (;;) {
                        visit(cur);
                        int d = 0;
                        for (; d < D; d++) {
                            if (++cur[d] <= hi[d]) break;
                            cur[d] = lo[d];
                        }
                        if (d == D) break;
                    }
                }
            }
        }
    }
    // boundary tie: a candidate left out at the m-th distance
    P.flag[k] = (len == m && len > 0 && rej == H.dd(0)) ? 1 : 0;
    // ascending (d2, position): heap sort in place
    for (int e = len - 1; e > 0; e--) {
        const double td = H.dd(0);
        const int32_t tp = H.pp(0);
        H.dd(0) = H.dd(e);
        H.pp(0) = H.pp(e);
        H.dd(e) = td;
        H.pp(e) = tp;
        H.sift_down(e, 0);
    }
    for (int e = 0; e < m; e++) P.out[k * m + e] = e < len ? H.pp(e) : -1;
    P.count[k] = len;
}

}  // namespace

hipError_t launch_pred_knn(int D, const PredParams &p, hipStream_t s) {
    const unsigned blocks = (unsigned)((p.n + kL - 1) / kL);
    const size_t lds = (sizeof(double) + sizeof(int32_t)) * (size_t)p.m * kL;
    if (D == 2) hipLaunchKernelGGL((pred_knn_kernel<2>), dim3(blocks), dim3(kL), lds, s, p);
    else if (D == 3) hipLaunchKernelGGL((pred_knn_kernel<3>), dim3(blocks), dim3(kL), lds, s, p);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace gs
