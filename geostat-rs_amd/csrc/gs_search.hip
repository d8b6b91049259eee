// Neighbour search on gfx950: exact emulation of the reference kd-tree query.
//
// KdTree::k_nearest (search.rs:57-99) is a depth-first median-split
// traversal with a bounded max-heap keyed by (d2 bits, idx): a full heap
// accepts a candidate only if d2 < worst (strict, on d2 alone) and then
// evicts the largest (d2, idx).  When the k-th and (k+1)-th distances tie,
// which index survives depends on the visiting order, so this kernel
// replays the reference traversal exactly (same tree layout, same near/far
// rule `diff <= 0 -> left`, same prune `diff^2 <= min(worst, r2)`) rather
// than computing a canonical k-smallest set.  Distances are
// ((0 + dx^2) + dy^2) [+ dz^2] with contraction disabled, as rustc computes
// them.
//
// One thread per target; targets arrive in grid order, so neighbouring
// lanes walk nearly the same tree path.  The heap lives in LDS with a
// lane-interleaved layout (entry e of lane l at [e][l]), conflict-free.
// The traversal stack (pending far subtrees) is a small private array.
#include <algorithm>

#include "gs_internal.hpp"

#pragma clang fp contract(off)

namespace gs {

template <int D>
__device__ __forceinline__ void search_xf(const SearchXf &X, const double *p, double *out) {
    if (!X.on) {
#pragma unroll
        for (int d = 0; d < D; d++) out[d] = p[d];
        return;
    }
    if (D == 2) {
        out[0] = p[0] * X.s + p[1] * X.c;
        out[1] = (p[0] * X.c - p[1] * X.s) / X.ratio;
    } else {
        out[0] = X.r[0] * p[0] + X.r[1] * p[1] + X.r[2] * p[2];
        out[1] = X.r[3] * p[0] + X.r[4] * p[1] + X.r[5] * p[2];
        out[2] = X.r[6] * p[0] + X.r[7] * p[1] + X.r[8] * p[2];
    }
}

constexpr int kStack = 48;   // kd-tree depth <= 33 for n < 2^32
constexpr int kBlock = 64;

// Max-heap in LDS over (d2, idx), lane-interleaved.
struct LHeap {
    double *d;
    uint32_t *ix;
    int lane;
    __device__ __forceinline__ double D2(int e) const { return d[e * kBlock + lane]; }
    __device__ __forceinline__ uint32_t IX(int e) const { return ix[e * kBlock + lane]; }
    __device__ __forceinline__ void set(int e, double v, uint32_t i) {
        d[e * kBlock + lane] = v;
        ix[e * kBlock + lane] = i;
    }
    static __device__ __forceinline__ bool gt(double ad, uint32_t ai, double bd, uint32_t bi) {
        return ad > bd || (ad == bd && ai > bi);
    }
    __device__ __forceinline__ void push(int &len, double v, uint32_t i) {
        int e = len++;
        while (e > 0) {
            int p = (e - 1) >> 1;
            double pd = D2(p);
            uint32_t pi = IX(p);
            if (gt(v, i, pd, pi)) { set(e, pd, pi); e = p; }
            else break;
        }
        set(e, v, i);
    }
    __device__ __forceinline__ void sift_down(int len, int e, double v, uint32_t i) {
        for (;;) {
            int l = 2 * e + 1;
            if (l >= len) break;
            int r = l + 1;
            int c = l;
            double cd = D2(l);
            uint32_t ci = IX(l);
            if (r < len) {
                double rd = D2(r);
                uint32_t ri = IX(r);
                if (gt(rd, ri, cd, ci)) { c = r; cd = rd; ci = ri; }
            }
            if (gt(cd, ci, v, i)) { set(e, cd, ci); e = c; }
            else break;
        }
        set(e, v, i);
    }
    // Heapsort in place: ascending (d2, idx), as into_sorted (search.rs:32-36).
    __device__ __forceinline__ void sort(int len) {
        for (int end = len - 1; end > 0; end--) {
            double td = D2(0);
            uint32_t ti = IX(0);
            double vd = D2(end);
            uint32_t vi = IX(end);
            set(end, td, ti);
            sift_down(end, 0, vd, vi);
        }
    }
};

template <int D>
__device__ __forceinline__ double dist2(const double *p, const double *t) {
    double d2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; d++) {
        double dd = p[d] - t[d];
        d2 += dd * dd;
    }
    return d2;
}

// Pending far subtree: slice [lo, hi) plus the parent's split (its position
// and axis) so the prune test can recompute diff^2 when the entry is popped.
struct StackEnt {
    uint32_t lo, hi;
};

template <int D, bool OCT>
__global__ __launch_bounds__(kBlock) void knn_kernel(KnnParams P) {
    extern __shared__ unsigned char smem_raw[];
    const int lane = threadIdx.x;
    const int slots = OCT ? (P.per_octant << D) : P.k;
    LHeap H;
    if (P.gheap_d) {   // large neighbourhoods: this block's heaps in device memory
        H.d = P.gheap_d + (size_t)blockIdx.x * slots * kBlock;
        H.ix = P.gheap_ix + (size_t)blockIdx.x * slots * kBlock;
    } else {
        H.d = reinterpret_cast<double *>(smem_raw);
        H.ix = reinterpret_cast<uint32_t *>(H.d + (size_t)slots * kBlock);
    }
    H.lane = lane;

    const int64_t nt = P.list && P.list_len ? (int64_t)(*P.list_len) : P.T;
    for (int64_t it = (int64_t)blockIdx.x * kBlock + lane; it < nt;
         it += (int64_t)gridDim.x * kBlock) {
    const int64_t t = P.list ? (int64_t)P.list[it] : it;
    double tg[D], st[D];
#pragma unroll
    for (int d = 0; d < D; d++) tg[d] = P.targets[t * D + d];
    search_xf<D>(P.xf, tg, st);

    const double r2 = P.r2;
    const int k = P.k;
    const int q = P.per_octant;
    int len = 0;                  // heap mode
    int olen[1 << D];             // octant mode: per-octant heap sizes
#pragma unroll
    for (int o = 0; o < (1 << D); o++) olen[o] = 0;

    StackEnt stk[kStack];
    uint32_t stk_meta[kStack];    // bit 31: far is the right child; bits 0..1: parent axis
    int sp = 0;
    uint32_t lo = 0, hi = (uint32_t)P.n;
    int axis = 0;
    const double *tc = P.tcoords;

    for (;;) {
        while (lo < hi) {
            uint32_t mid = lo + ((hi - lo) >> 1);
            const double *p = tc + (size_t)mid * D;
            double pc[D];
#pragma unroll
            for (int d = 0; d < D; d++) pc[d] = p[d];
            double d2 = dist2<D>(pc, st);
            if (d2 <= r2) {
                uint32_t id = P.tidx[mid];
                if (!OCT) {
                    // push_candidate (search.rs:15-22)
                    if (len < k) H.push(len, d2, id);
                    else if (d2 < H.D2(0)) H.sift_down(len, 0, d2, id);
                } else {
                    // Octant mode queries k = n, so the reference heap never
                    // saturates and the candidate list is canonical (all
                    // in-radius points by (d2, idx)); keep the per-octant
                    // first `q` of that order (kriging.rs:471-485).
                    int oct = 0;
#pragma unroll
                    for (int d = 0; d < D; d++)
                        if (pc[d] >= st[d]) oct |= 1 << d;
                    LHeap O = H;
                    O.d = H.d + (size_t)oct * q * kBlock;
                    O.ix = H.ix + (size_t)oct * q * kBlock;
                    int ol = olen[0];
#pragma unroll
                    for (int o = 1; o < (1 << D); o++) if (oct == o) ol = olen[o];
                    if (ol < q) {
                        O.push(ol, d2, id);
#pragma unroll
                        for (int o = 0; o < (1 << D); o++) if (oct == o) olen[o] = ol;
                    } else if (LHeap::gt(O.D2(0), O.IX(0), d2, id)) {
                        O.sift_down(ol, 0, d2, id);
                    }
                }
            }
            double diff = st[axis] - pc[axis];
            uint32_t nlo, nhi, flo, fhi, right;
            if (diff <= 0.0) { nlo = lo; nhi = mid; flo = mid + 1; fhi = hi; right = 1u; }
            else { nlo = mid + 1; nhi = hi; flo = lo; fhi = mid; right = 0u; }
            if (flo < fhi && sp < kStack) {
                stk[sp].lo = flo;
                stk[sp].hi = fhi;
                stk_meta[sp] = (right << 31) | (uint32_t)axis;
                sp++;
            }
            lo = nlo;
            hi = nhi;
            axis = (axis + 1 == D) ? 0 : axis + 1;
        }
        bool found = false;
        while (sp > 0) {
            sp--;
            uint32_t flo = stk[sp].lo, fhi = stk[sp].hi, meta = stk_meta[sp];
            int paxis = (int)(meta & 3u);
            uint32_t pmid = (meta >> 31) ? flo - 1 : fhi;
            double diff = st[paxis] - tc[(size_t)pmid * D + paxis];
            double worst = INFINITY;
            if (!OCT && len == k) worst = H.D2(0);
            if (diff * diff <= fmin(worst, r2)) {
                lo = flo;
                hi = fhi;
                axis = (paxis + 1 == D) ? 0 : paxis + 1;
                found = true;
                break;
            }
        }
        if (!found) break;
    }

    const int64_t oi = P.out_by_pos ? it : t;
    int32_t *out = P.out_idx + oi * P.stride;
    if (!OCT) {
        H.sort(len);
        for (int e = 0; e < len; e++) out[e] = (int32_t)H.IX(e);
        for (int e = len; e < P.stride; e++) out[e] = -1;
        P.out_count[oi] = len;
        continue;
    }
    // Octant mode: sort each octant, then merge by (d2, idx) and truncate
    // to max_neighbors (kriging.rs:486-488).
    int head[1 << D];
#pragma unroll
    for (int o = 0; o < (1 << D); o++) {
        LHeap O = H;
        O.d = H.d + (size_t)o * q * kBlock;
        O.ix = H.ix + (size_t)o * q * kBlock;
        O.sort(olen[o]);
        head[o] = 0;
    }
    int total = 0;
#pragma unroll
    for (int o = 0; o < (1 << D); o++) total += olen[o];
    int outn = total < P.max_nb ? total : P.max_nb;
    for (int e = 0; e < outn; e++) {
        int best = -1;
        double bd = 0.0;
        uint32_t bi = 0;
#pragma unroll
        for (int o = 0; o < (1 << D); o++) {
            if (head[o] < olen[o]) {
                double od = H.d[((size_t)o * q + head[o]) * kBlock + lane];
                uint32_t oi = H.ix[((size_t)o * q + head[o]) * kBlock + lane];
                if (best < 0 || LHeap::gt(bd, bi, od, oi)) { best = o; bd = od; bi = oi; }
            }
        }
#pragma unroll
        for (int o = 0; o < (1 << D); o++) if (best == o) head[o]++;
        out[e] = (int32_t)bi;
    }
    for (int e = outn; e < P.stride; e++) out[e] = -1;
    P.out_count[oi] = outn;
    }
}

// Number of data points within the search radius of each listed target:
// the kd-tree walk of knn_kernel without a heap (every subtree whose split
// plane is within the radius is visited).
template <int D>
__global__ __launch_bounds__(kBlock) void radius_count_kernel(KnnParams P, int32_t *counts) {
    const int64_t nt = P.list && P.list_len ? (int64_t)(*P.list_len) : P.T;
    for (int64_t it = (int64_t)blockIdx.x * kBlock + threadIdx.x; it < nt;
         it += (int64_t)gridDim.x * kBlock) {
        const int64_t t = P.list ? (int64_t)P.list[it] : it;
        double tg[D], st[D];
#pragma unroll
        for (int d = 0; d < D; d++) tg[d] = P.targets[t * D + d];
        search_xf<D>(P.xf, tg, st);
        const double r2 = P.r2;
        const double *tc = P.tcoords;
        StackEnt stk[kStack];
        uint32_t stk_axis[kStack];
        int sp = 0;
        uint32_t lo = 0, hi = (uint32_t)P.n;
        int axis = 0;
        int32_t cnt = 0;
        for (;;) {
            while (lo < hi) {
                const uint32_t mid = lo + ((hi - lo) >> 1);
                double pc[D];
#pragma unroll
                for (int d = 0; d < D; d++) pc[d] = tc[(size_t)mid * D + d];
                if (dist2<D>(pc, st) <= r2) cnt++;
                const double diff = st[axis] - pc[axis];
                uint32_t nlo, nhi, flo, fhi;
                if (diff <= 0.0) { nlo = lo; nhi = mid; flo = mid + 1; fhi = hi; }
                else { nlo = mid + 1; nhi = hi; flo = lo; fhi = mid; }
                if (flo < fhi && diff * diff <= r2 && sp < kStack) {
                    stk[sp].lo = flo;
                    stk[sp].hi = fhi;
                    stk_axis[sp] = (uint32_t)axis;
                    sp++;
                }
                lo = nlo;
                hi = nhi;
                axis = (axis + 1 == D) ? 0 : axis + 1;
            }
            if (sp == 0) break;
            sp--;
            lo = stk[sp].lo;
            hi = stk[sp].hi;
            axis = ((int)stk_axis[sp] + 1 == D) ? 0 : (int)stk_axis[sp] + 1;
        }
        counts[it] = cnt;
    }
}

hipError_t launch_radius_count(int D, const KnnParams &p, int32_t *counts, hipStream_t s) {
    if (p.T == 0) return hipSuccess;
    int64_t blocks = (p.T + kBlock - 1) / kBlock;
    if (p.list) blocks = std::min<int64_t>(blocks, 2048);
    if (D == 2) hipLaunchKernelGGL((radius_count_kernel<2>), dim3((unsigned)blocks), dim3(kBlock), 0, s, p, counts);
    else if (D == 3) hipLaunchKernelGGL((radius_count_kernel<3>), dim3((unsigned)blocks), dim3(kBlock), 0, s, p, counts);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_knn(int D, const KnnParams &p, hipStream_t s) {
    if (p.T == 0) return hipSuccess;
    const bool oct = p.per_octant > 0;
    const int slots = oct ? (p.per_octant << D) : p.k;
    const size_t lds = p.gheap_d ? 0 : (size_t)slots * kBlock * (sizeof(double) + sizeof(uint32_t));
    int64_t blocks = (p.T + kBlock - 1) / kBlock;
    if (p.list) blocks = std::min<int64_t>(blocks, 2048);   // grid-stride over the list
    if (p.gheap_d) blocks = std::min<int64_t>(blocks, p.heap_blocks);
    const dim3 grid((unsigned)blocks);
    if (D == 2) {
        if (oct) hipLaunchKernelGGL((knn_kernel<2, true>), grid, dim3(kBlock), lds, s, p);
        else hipLaunchKernelGGL((knn_kernel<2, false>), grid, dim3(kBlock), lds, s, p);
    } else if (D == 3) {
        if (oct) hipLaunchKernelGGL((knn_kernel<3, true>), grid, dim3(kBlock), lds, s, p);
        else hipLaunchKernelGGL((knn_kernel<3, false>), grid, dim3(kBlock), lds, s, p);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace gs
