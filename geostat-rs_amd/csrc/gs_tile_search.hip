// Tile neighbour search on gfx950: one wavefront per 64 spatially close
// targets, walking cell rings of a uniform grid in lockstep.
//
// Result contract: the same lists as KdTree::k_nearest (search.rs:57-99).
// The reference keeps a bounded max-heap keyed (d2 bits, idx) that accepts a
// candidate into a full heap only if d2 < worst; its final set is every point
// with d2 below the k-th distance W plus some points with d2 == W, chosen by
// the traversal order.  When no excluded point has d2 == W that set is the
// canonical "k smallest by (d2, idx)", which is what each lane computes here
// (visiting order does not matter for a canonical selection).  A lane also
// records the smallest distance it rejected or evicted; if that equals W the
// target is a boundary tie and is handed to the exact kd-tree replay
// (gs_search.hip), which reproduces the reference's traversal order.
// Octant lists (kriging.rs:471-488 with k = n) are canonical by construction.
//
// Distances use the reference's expression ((0 + dx^2) + dy^2) [+ dz^2] with
// contraction off, on search-frame coordinates transformed exactly as the
// reference does, so every compared d2 is bit-identical.
//
// Ordering: targets are sorted by the Hilbert index of their search-frame
// cell (Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707,
// 2004: axes-to-transpose, then bit interleave), so consecutive targets lie
// in adjacent cells and a wavefront covers a small box of cells.  A
// wavefront whose box exceeds kMaxBoxCells (the curve leaving and
// re-entering a non-power-of-two grid, or sparse targets) hands all its
// targets to the kd-tree replay instead.
//
// Walk: ring r is the box grown by r cells (clipped to the grid) minus the
// region already visited, enumerated as row segments (x-runs).  The 64 lanes
// resolve up to 64 rows' segments at once, prefix-sum their lengths, and
// then load up to 64 candidates per step cooperatively; each candidate is
// broadcast with v_readlane and offered to every lane's top-k.  After each
// ring a lane stops when the distance from its target to the unvisited part
// of the grid (outside the visited rectangle) exceeds its k-th distance or
// the search radius; the wavefront stops when all lanes have.

#include <hipcub/hipcub.hpp>

#include "gs_internal.hpp"

#pragma clang fp contract(off)

namespace gs {

namespace {

constexpr int kTile = 64;
constexpr int kMaxBoxCells = 64;

template <int D>
__device__ __forceinline__ void xf_target(const SearchXf &X, const double *p, double *out) {
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

// Cell of a search-frame position, clamped to the grid (same rule as the
// host's point binning); NaN maps to cell 0.
template <int D>
__device__ __forceinline__ void cell_of(const CellGridView &G, const double *p, int *c) {
#pragma unroll
    for (int d = 0; d < D; d++) {
        const double f = (p[d] - G.origin[d]) * G.inv_cell;
        int v;
        if (!(f == f)) v = 0;
        else if (f <= 0.0) v = 0;
        else if (f >= (double)(G.nc[d] - 1)) v = G.nc[d] - 1;
        else v = (int)f;
        c[d] = v;
    }
}

// Hilbert index of a cell, `bits` bits per axis (Skilling 2004).
template <int D>
__device__ __forceinline__ uint64_t hilbert_index(const int *c, int bits) {
    uint32_t X[D];
#pragma unroll
    for (int i = 0; i < D; i++) X[i] = (uint32_t)c[i];
    const uint32_t M = 1u << (bits - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
#pragma unroll
        for (int i = 0; i < D; i++) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
#pragma unroll
    for (int i = 1; i < D; i++) X[i] ^= X[i - 1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[D - 1] & Q) t ^= Q - 1;
#pragma unroll
    for (int i = 0; i < D; i++) X[i] ^= t;
    uint64_t key = 0;
    for (int b = bits - 1; b >= 0; b--) {
#pragma unroll
        for (int i = 0; i < D; i++) key = (key << 1) | ((X[i] >> b) & 1u);
    }
    return key;
}

// Hilbert key of every grid cell (cell id = x + nc0 * (y + nc1 * z)), with
// the cell id as the sort payload; `shift` coarsens the cells when the
// interleaved index would not fit in 64 bits.
template <int D>
__global__ void cell_key_kernel(CellGridView G, int64_t ncells, int bits, int shift,
                                uint64_t *keys, uint32_t *ids) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= ncells) return;
    int c[D];
    int64_t r = id;
#pragma unroll
    for (int d = 0; d < D; d++) {
        c[d] = (int)(r % G.nc[d]) >> shift;
        r /= G.nc[d];
    }
    keys[id] = hilbert_index<D>(c, bits);
    ids[id] = (uint32_t)id;
}

__global__ void rank_scatter_kernel(const uint32_t *sorted_ids, int64_t ncells, uint32_t *rank) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ncells) rank[sorted_ids[i]] = (uint32_t)i;
}

// Per target: the Hilbert rank of its search-frame cell, >> bshift.
template <int D>
__global__ void target_bin_kernel(const double *targets, int64_t T, SearchXf xf, CellGridView G,
                                  const uint32_t *rank, int bshift, uint32_t *bins) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    double p[D], st[D];
#pragma unroll
    for (int d = 0; d < D; d++) p[d] = targets[t * D + d];
    xf_target<D>(xf, p, st);
    int c[D];
    cell_of<D>(G, st, c);
    int64_t id = 0, mul = 1;
#pragma unroll
    for (int d = 0; d < D; d++) {
        id += (int64_t)c[d] * mul;
        mul *= G.nc[d];
    }
    bins[t] = rank[id] >> bshift;
}

// Bounded max-heap over (d2, idx) in LDS, lane-interleaved (entry e of
// lane l at [e][l]: conflict-free).
// Bounded heap per lane, lane-interleaved in LDS.  A slot holds the
// candidate's squared distance rounded to float (the ordering key) and its
// position in the cell-ordered point arrays: 8 bytes instead of 12, so more
// wavefronts fit in LDS.  Rounding is monotone, so the float order agrees
// with the exact (d2, index) order except between candidates whose keys
// round to the same float: the kernel flags a target when such a tie sits
// at the selection boundary (the exact kd-tree replay then decides) and
// re-orders equal-key runs of the selected set by the exact (d2, index).
struct TileHeap {
    float *d;
    uint32_t *ix;
    int lane;
    __device__ __forceinline__ float D2(int e) const { return d[e * kTile + lane]; }
    __device__ __forceinline__ uint32_t IX(int e) const { return ix[e * kTile + lane]; }
    __device__ __forceinline__ void set(int e, float v, uint32_t i) {
        d[e * kTile + lane] = v;
        ix[e * kTile + lane] = i;
    }
    static __device__ __forceinline__ bool gt(float ad, uint32_t ai, float bd, uint32_t bi) {
        return ad > bd || (ad == bd && ai > bi);
    }
    __device__ __forceinline__ void push(int &len, float v, uint32_t i) {
        int e = len++;
        while (e > 0) {
            const int p = (e - 1) >> 1;
            const float pd = D2(p);
            const uint32_t pi = IX(p);
            if (!gt(v, i, pd, pi)) break;
            set(e, pd, pi);
            e = p;
        }
        set(e, v, i);
    }
    // Sift (v, i) down from slot e; returns the value left in slot e.
    __device__ __forceinline__ void sift_down(int len, int e, float v, uint32_t i, float &top_d,
                                              uint32_t &top_i) {
        top_d = v;
        top_i = i;
        bool first = true;
        for (;;) {
            const int l = 2 * e + 1;
            if (l >= len) break;
            int c = l;
            float cd = D2(l);
            uint32_t ci = IX(l);
            if (l + 1 < len) {
                const float rd = D2(l + 1);
                const uint32_t ri = IX(l + 1);
                if (gt(rd, ri, cd, ci)) { c = l + 1; cd = rd; ci = ri; }
            }
            if (!gt(cd, ci, v, i)) break;
            set(e, cd, ci);
            if (first) { top_d = cd; top_i = ci; first = false; }
            e = c;
        }
        set(e, v, i);
    }
    // Floyd's bottom-up construction over slots [0, len).
    __device__ __forceinline__ void heapify(int len) {
        float td;
        uint32_t ti;
        for (int e = len / 2 - 1; e >= 0; e--) sift_down(len, e, D2(e), IX(e), td, ti);
    }
    __device__ __forceinline__ void sort(int len) {
        float td;
        uint32_t ti;
        for (int end = len - 1; end > 0; end--) {
            const float rd = D2(0);
            const uint32_t ri = IX(0);
            const float vd = D2(end);
            const uint32_t vi = IX(end);
            set(end, rd, ri);
            sift_down(end, 0, vd, vi, td, ti);
        }
    }
};

// Ascending bitonic sorting network over the first N entries of a register
// array (N a power of two; indices are compile-time constants, so the array
// stays in VGPRs).  Used for the final order of small selections, where a
// heap sort in LDS would be a chain of dependent LDS round trips.
template <int N, int CAP>
__device__ __forceinline__ void bitonic_sort(uint64_t (&a)[CAP]) {
    static_assert(N <= CAP && (N & (N - 1)) == 0, "bitonic size");
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
            for (int i = 0; i < N; i++) {
                const int l = i ^ j;
                if (l > i) {
                    const uint64_t x = a[i], y = a[l];
                    const bool up = (i & k) == 0;
                    const bool sw = up ? x > y : x < y;
                    a[i] = sw ? y : x;
                    a[l] = sw ? x : y;
                }
            }
        }
    }
}

// Upper bound of the exact squared distance behind a float key (round to
// nearest: within half an ulp).
__device__ __forceinline__ double key_upper(float f) { return (double)f * (1.0 + 1.2e-7); }

// Exact search-frame squared distance of the point at cell-order position
// p, in the reference's operation order.
template <int D>
__device__ __forceinline__ double exact_d2(const CellGridView &G, uint32_t p, const double *st) {
    double d2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; d++) {
        const double dd = G.pts[(int64_t)p * D + d] - st[d];
        d2 += dd * dd;
    }
    return d2;
}

// (d2, original index) order of two cell-order positions.
template <int D>
__device__ __forceinline__ bool exact_gt(const CellGridView &G, uint32_t a, uint32_t b,
                                         const double *st) {
    const double da = exact_d2<D>(G, a, st), db = exact_d2<D>(G, b, st);
    return da > db || (da == db && G.idx[a] > G.idx[b]);
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
#pragma unroll
    for (int o = 1; o < kTile; o <<= 1) {
        const int y = __shfl_up(v, o);
        if (lane >= o) v += y;
    }
    return v;
}
__device__ __forceinline__ double readlane_f64(double v, int j) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, j);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), j);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Dynamic LDS: the heaps ([slots][64] doubles, then [slots][64] indices),
// then the row table of the ring being walked.
struct RowTable {
    int32_t excl[kTile];   // exclusive prefix of the rows' candidate counts
    int32_t alen[kTile];   // length of the row's first segment
    int32_t a0[kTile];     // first point of segment a
    int32_t b0[kTile];     // first point of segment b
};

template <int D, bool OCT>
__global__ __launch_bounds__(kTile) __attribute__((amdgpu_waves_per_eu(3))) void tile_knn_kernel(TileParams P) {
    extern __shared__ unsigned char smem_raw[];
    const int lane = threadIdx.x;
    const int slots = OCT ? (P.per_octant << D) : P.k;
    TileHeap H;
    H.d = reinterpret_cast<float *>(smem_raw);
    H.ix = reinterpret_cast<uint32_t *>(H.d + (size_t)slots * kTile);
    H.lane = lane;
    RowTable *RT = reinterpret_cast<RowTable *>(smem_raw + (size_t)slots * kTile * 8);
    const CellGridView &G = P.G;

    const int64_t pos = xcd_block(blockIdx.x, gridDim.x) * kTile + lane;
    const bool live = pos < P.T;
    const int64_t t = live ? (int64_t)P.perm[pos] : -1;
    double st[D], ft[D];
    int cl[3] = {0, 0, 0};
    bool nan_target = false;
    if (live) {
        double p[D];
#pragma unroll
        for (int d = 0; d < D; d++) p[d] = P.targets[t * D + d];
        xf_target<D>(P.xf, p, st);
#pragma unroll
        for (int d = 0; d < D; d++) {
            nan_target = nan_target || !(st[d] == st[d]);
            ft[d] = (st[d] - G.origin[d]) * G.inv_cell;
        }
        cell_of<D>(G, st, cl);
    } else {
#pragma unroll
        for (int d = 0; d < D; d++) st[d] = ft[d] = 0.0;
    }
    const bool active = live && !nan_target;

    // box of the active lanes' cells
    const uint64_t active_mask = __ballot(active);
    int blo[3] = {0, 0, 0}, bhi[3] = {0, 0, 0};
    int64_t box_cells = 0;
    if (active_mask != 0) {
        box_cells = 1;
#pragma unroll
        for (int d = 0; d < D; d++) {
            blo[d] = wave_min(active ? cl[d] : INT32_MAX);
            bhi[d] = wave_max(active ? cl[d] : INT32_MIN);
            box_cells *= (int64_t)(bhi[d] - blo[d] + 1);
        }
    }
    if (active_mask != 0 && box_cells > kMaxBoxCells) {
        // scattered wavefront: every live target goes to the kd-tree replay
        const uint64_t live_mask = __ballot(live);
        int base = 0;
        if (lane == 0) base = atomicAdd(P.flag_len, (int)__popcll(live_mask));
        base = __builtin_amdgcn_readfirstlane(base);
        if (live) {
            const int rank = (int)__popcll(live_mask & ((1ull << lane) - 1ull));
            P.flag_list[base + rank] = (int32_t)t;
            if (P.flag_mark) P.flag_mark[t] = 1;
        }
        return;
    }

    const double r2 = P.r2;
    const int k = P.k;
    const int q = P.per_octant;
    int len = 0;
    float top_d = INFINITY;         // heap mode: root of the full heap
    uint32_t top_i = 0;
    int olen[1 << D];
    float orej[1 << D];             // octant mode: smallest key left out, per octant
#pragma unroll
    for (int o = 0; o < (1 << D); o++) {
        olen[o] = 0;
        orej[o] = INFINITY;
    }
    float rej = INFINITY;
    float wmax = INFINITY;          // octant mode: largest root once all heaps are full
    bool done = !active;

    int vlo[3] = {0, 0, 0}, vhi[3] = {0, 0, 0};   // visited rectangle
    bool visited_any = false;
    for (int r = 0; active_mask != 0; r++) {
        int elo[3] = {0, 0, 0}, ehi[3] = {0, 0, 0};
        bool grew = !visited_any;
#pragma unroll
        for (int d = 0; d < D; d++) {
            elo[d] = max(blo[d] - r, 0);
            ehi[d] = min(bhi[d] + r, G.nc[d] - 1);
            grew = grew || elo[d] != vlo[d] || ehi[d] != vhi[d];
        }
        if (!grew) break;                           // whole grid visited
        const int ny = D >= 2 ? ehi[1] - elo[1] + 1 : 1;
        const int nz = D == 3 ? ehi[2] - elo[2] + 1 : 1;
        const int nrows = ny * nz;
        for (int rb = 0; rb < nrows; rb += kTile) {
            const int ri = rb + lane;
            int32_t a0 = 0, alen = 0, b0 = 0, blen = 0;
            if (ri < nrows) {
                const int y = D >= 2 ? elo[1] + ri % ny : 0;
                const int z = D == 3 ? elo[2] + ri / ny : 0;
                bool inner = visited_any;
                if (D >= 2) inner = inner && y >= vlo[1] && y <= vhi[1];
                if (D == 3) inner = inner && z >= vlo[2] && z <= vhi[2];
                const int64_t row = ((int64_t)z * G.nc[1] + y) * G.nc[0];
                if (!inner) {
                    a0 = G.start[row + elo[0]];
                    alen = G.start[row + ehi[0] + 1] - a0;
                } else {
                    if (elo[0] < vlo[0]) {
                        a0 = G.start[row + elo[0]];
                        alen = G.start[row + elo[0] + 1] - a0;
                    }
                    if (ehi[0] > vhi[0]) {
                        b0 = G.start[row + ehi[0]];
                        blen = G.start[row + ehi[0] + 1] - b0;
                    }
                }
            }
            const int rlen = alen + blen;
            const int incl = wave_incl_scan(rlen, lane);
            const int total = __builtin_amdgcn_readlane(incl, kTile - 1);
            RT->excl[lane] = incl - rlen;
            RT->alen[lane] = alen;
            RT->a0[lane] = a0;
            RT->b0[lane] = b0;
            __syncthreads();
            for (int cb = 0; cb < total; cb += kTile) {
                if (OCT) {
                    // once every octant heap is full, a candidate whose key
                    // exceeds all their roots enters none of them (the bound
                    // is refreshed per chunk: roots only fall in between)
                    bool full = true;
                    float m = 0.0f;
#pragma unroll
                    for (int o = 0; o < (1 << D); o++) {
                        full = full && olen[o] == q;
                        if (olen[o] > 0) m = fmaxf(m, H.d[((size_t)o * q) * kTile + lane]);
                    }
                    wmax = full ? m : INFINITY;
                }
                const int g = cb + lane;
                double pc[D];
#pragma unroll
                for (int d = 0; d < D; d++) pc[d] = 0.0;
                uint32_t pid = 0;   // cell-order position
                if (g < total) {
                    int at = 0;    // last row whose exclusive prefix is <= g
#pragma unroll
                    for (int step = kTile / 2; step > 0; step >>= 1)
                        if (RT->excl[at + step] <= g) at += step;
                    const int o = g - RT->excl[at];
                    const int al = RT->alen[at];
                    const int64_t p = o < al ? (int64_t)RT->a0[at] + o : (int64_t)RT->b0[at] + (o - al);
#pragma unroll
                    for (int d = 0; d < D; d++) pc[d] = G.pts[p * D + d];
                    pid = (uint32_t)p;
                }
                const int cnt = min(kTile, total - cb);
                for (int j = 0; j < cnt; j++) {
                    double cp[D];
#pragma unroll
                    for (int d = 0; d < D; d++) cp[d] = readlane_f64(pc[d], j);
                    const uint32_t id = (uint32_t)__builtin_amdgcn_readlane((int)pid, j);
                    if (done) continue;
                    double d2 = 0.0;
#pragma unroll
                    for (int d = 0; d < D; d++) {
                        const double dd = cp[d] - st[d];
                        d2 += dd * dd;
                    }
                    if (!(d2 <= r2)) continue;
                    const float kf = __double2float_rn(d2);
                    if (OCT && kf > wmax) continue;
                    if (!OCT) {
                        if (len < k) {
                            // fill unordered; one bottom-up heapify when full
                            // (candidates arrive roughly by increasing
                            // distance, so sift-up pushes would climb to the
                            // root every time)
                            H.set(len, kf, id);
                            if (++len == k) {
                                H.heapify(k);
                                top_d = H.D2(0);
                                top_i = H.IX(0);
                            }
                        } else if (TileHeap::gt(top_d, top_i, kf, id)) {
                            rej = fminf(rej, top_d);
                            H.sift_down(len, 0, kf, id, top_d, top_i);
                        } else {
                            rej = fminf(rej, kf);
                        }
                    } else {
                        int oct = 0;
#pragma unroll
                        for (int d = 0; d < D; d++)
                            if (cp[d] >= st[d]) oct |= 1 << d;
                        TileHeap O = H;
                        O.d = H.d + (size_t)oct * q * kTile;
                        O.ix = H.ix + (size_t)oct * q * kTile;
                        int ol = olen[0];
#pragma unroll
                        for (int o = 1; o < (1 << D); o++) if (oct == o) ol = olen[o];
                        if (ol < q) {
                            O.push(ol, kf, id);
#pragma unroll
                            for (int o = 0; o < (1 << D); o++) if (oct == o) olen[o] = ol;
                        } else {
                            const float od = O.D2(0);
                            float left = kf;
                            if (TileHeap::gt(od, O.IX(0), kf, id)) {
                                float td;
                                uint32_t ti;
                                O.sift_down(ol, 0, kf, id, td, ti);
                                left = od;
                            }
#pragma unroll
                            for (int o = 0; o < (1 << D); o++)
                                if (oct == o) orej[o] = fminf(orej[o], left);
                        }
                    }
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int d = 0; d < D; d++) { vlo[d] = elo[d]; vhi[d] = ehi[d]; }
        visited_any = true;

        if (!done) {
            // distance (in cells) from the target to the unvisited cells:
            // points of a cell j have j <= f <= j + 1 on that axis
            double lbc = INFINITY;
#pragma unroll
            for (int d = 0; d < D; d++) {
                if (vlo[d] > 0) lbc = fmin(lbc, ft[d] - (double)vlo[d]);
                if (vhi[d] < G.nc[d] - 1) lbc = fmin(lbc, (double)(vhi[d] + 1) - ft[d]);
            }
            if (lbc == INFINITY) {
                done = true;                        // nothing left unvisited
            } else {
                // slack for the rounding of f and of the compared d2
                lbc = lbc * (1.0 - 1e-9) - 1e-6;
                double bound = r2;
                if (!OCT) {
                    if (len == k) bound = fmin(bound, key_upper(top_d));
                } else {
                    bool full = true;
                    double worst = 0.0;
#pragma unroll
                    for (int o = 0; o < (1 << D); o++) {
                        full = full && olen[o] == q;
                        if (olen[o] > 0)
                            worst = fmax(worst, key_upper(H.d[((size_t)o * q) * kTile + lane]));
                    }
                    if (full) bound = fmin(bound, worst);
                }
                if (lbc > 0.0) {
                    const double lb = lbc * G.cell;
                    if (lb * lb > bound) done = true;
                }
            }
        }
        if (__ballot(!done) == 0) break;
    }
    if (!live) return;

    int32_t *out = P.out_idx + t * P.stride;
    auto flag_target = [&]() {
        const int slot = atomicAdd(P.flag_len, 1);
        P.flag_list[slot] = (int32_t)t;
        if (P.flag_mark) P.flag_mark[t] = 1;
    };
    if (!OCT) {
        // a candidate left out with the boundary key: the exact order decides
        const bool tie = len == k && rej == top_d;
        if (k <= 32) {
            // (key, position) packed so that one unsigned compare orders
            // them (keys are non-negative floats); empty slots sort last
            uint64_t a[32];
#pragma unroll
            for (int e = 0; e < 32; e++) {
                a[e] = ~0ull;
                if (e < len) a[e] = ((uint64_t)__float_as_uint(H.D2(e)) << 32) | H.IX(e);
            }
            if (k <= 16) bitonic_sort<16>(a);
            else bitonic_sort<32>(a);
            // adjacent equal keys (rare) need the exact order: back to LDS
            bool run = false;
#pragma unroll
            for (int e = 1; e < 32; e++)
                run = run || (e < len && (uint32_t)(a[e - 1] >> 32) == (uint32_t)(a[e] >> 32));
            if (!run) {
#pragma unroll
                for (int e = 0; e < 32; e++)
                    if (e < len) out[e] = (int32_t)G.idx[(uint32_t)a[e]];
                for (int e = len; e < P.stride; e++) out[e] = -1;
                P.out_count[t] = len;
                if (tie) flag_target();
                return;
            }
#pragma unroll
            for (int e = 0; e < 32; e++)
                if (e < len) H.set(e, __uint_as_float((uint32_t)(a[e] >> 32)), (uint32_t)a[e]);
        } else {
            if (len < k) H.heapify(len);     // a partial fill is still unordered
            H.sort(len);
        }
        // runs of equal keys in exact (d2, index) order (insertion sort)
        for (int e = 1; e < len; e++) {
            for (int j = e; j > 0 && H.D2(j - 1) == H.D2(j) &&
                            exact_gt<D>(G, H.IX(j - 1), H.IX(j), st); j--) {
                const uint32_t a = H.IX(j - 1);
                H.set(j - 1, H.D2(j), H.IX(j));
                H.set(j, H.D2(j), a);
            }
        }
        for (int e = 0; e < len; e++) out[e] = (int32_t)G.idx[H.IX(e)];
        for (int e = len; e < P.stride; e++) out[e] = -1;
        P.out_count[t] = len;
        if (tie) flag_target();
        return;
    }
    bool tie = false;
    int head[1 << D];
#pragma unroll
    for (int o = 0; o < (1 << D); o++) {
        TileHeap O = H;
        O.d = H.d + (size_t)o * q * kTile;
        O.ix = H.ix + (size_t)o * q * kTile;
        O.sort(olen[o]);
        // a full octant whose boundary key is shared by a left-out candidate
        if (olen[o] == q && q > 0 && orej[o] == O.D2(q - 1)) tie = true;
        head[o] = 0;
    }
    int total = 0;
#pragma unroll
    for (int o = 0; o < (1 << D); o++) total += olen[o];
    const int outn = total < P.max_nb ? total : P.max_nb;
    // merge by key; positions go to out first, then equal-key runs are put in
    // exact order and the positions mapped to indices
    float last = -1.0f;
    for (int e = 0; e < total; e++) {
        int best = -1;
        float bd = 0.0f;
        uint32_t bi = 0;
#pragma unroll
        for (int o = 0; o < (1 << D); o++) {
            if (head[o] < olen[o]) {
                const float od = H.d[((size_t)o * q + head[o]) * kTile + lane];
                const uint32_t oi = H.ix[((size_t)o * q + head[o]) * kTile + lane];
                if (best < 0 || TileHeap::gt(bd, bi, od, oi)) { best = o; bd = od; bi = oi; }
            }
        }
#pragma unroll
        for (int o = 0; o < (1 << D); o++) if (best == o) head[o]++;
        if (e < outn) {
            out[e] = (int32_t)bi;
            last = bd;
        } else {
            // truncation inside a run of equal keys: the exact order decides
            if (bd == last) tie = true;
            break;
        }
    }
    for (int e = 1; e < outn; e++) {
        for (int j = e; j > 0; j--) {
            const uint32_t a = (uint32_t)out[j - 1], b = (uint32_t)out[j];
            if (__double2float_rn(exact_d2<D>(G, a, st)) != __double2float_rn(exact_d2<D>(G, b, st)))
                break;
            if (!exact_gt<D>(G, a, b, st)) break;
            out[j - 1] = (int32_t)b;
            out[j] = (int32_t)a;
        }
    }
    for (int e = 0; e < outn; e++) out[e] = (int32_t)G.idx[(uint32_t)out[e]];
    for (int e = outn; e < P.stride; e++) out[e] = -1;
    P.out_count[t] = outn;
    if (tie) flag_target();
}

}  // namespace

// Hilbert rank of every cell of the grid (build time; allocates its own
// temporaries).  Consecutive ranks are adjacent cells, so targets taken in
// rank order share cells and neighbourhoods.
hipError_t build_cell_ranks(int D, const CellGridView &G, uint32_t *rank, hipStream_t s) {
    if (D != 2 && D != 3) return hipErrorInvalidValue;
    const int64_t ncells = (int64_t)G.nc[0] * G.nc[1] * G.nc[2];
    // bits per axis covering the largest cell index; coarser cells when the
    // interleaved index would not fit in 64 bits
    int need = 1;
    for (int d = 0; d < D; d++)
        while (need < 31 && (int64_t)(G.nc[d] - 1) >> need) need++;
    const int bits = std::min(need, 64 / D);
    const int shift = need - bits;
    uint64_t *keys = nullptr, *keys_out = nullptr;
    uint32_t *ids = nullptr, *ids_out = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    hipError_t e = hipMalloc(&keys, sizeof(uint64_t) * ncells);
    if (e == hipSuccess) e = hipMalloc(&keys_out, sizeof(uint64_t) * ncells);
    if (e == hipSuccess) e = hipMalloc(&ids, sizeof(uint32_t) * ncells);
    if (e == hipSuccess) e = hipMalloc(&ids_out, sizeof(uint32_t) * ncells);
    if (e == hipSuccess)
        e = hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys, keys_out, ids, ids_out,
                                               (int)ncells, 0, D * bits, s);
    if (e == hipSuccess) e = hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 8);
    const dim3 grid((unsigned)((ncells + 255) / 256));
    if (e == hipSuccess) {
        if (D == 2)
            hipLaunchKernelGGL(cell_key_kernel<2>, grid, dim3(256), 0, s, G, ncells, bits, shift,
                               keys, ids);
        else
            hipLaunchKernelGGL(cell_key_kernel<3>, grid, dim3(256), 0, s, G, ncells, bits, shift,
                               keys, ids);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, keys_out, ids, ids_out,
                                               (int)ncells, 0, D * bits, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(rank_scatter_kernel, grid, dim3(256), 0, s, ids_out, ncells, rank);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (void *p : {(void *)keys, (void *)keys_out, (void *)ids, (void *)ids_out, tmp})
        if (p) (void)hipFree(p);
    return e;
}

// Target order: a counting sort of the targets by the Hilbert rank of their
// cell -- histogram, exclusive scan, scatter.  Targets of one bin (one cell,
// or a few consecutive cells for calls with far fewer targets than cells)
// keep no particular order among themselves, which no result depends on:
// every neighbour list is canonical and every system is solved on its own.
// Only kernels, with no memset and no host synchronisation, so a predict
// call can be captured into a HIP graph.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanChunk = kScanThreads * kScanItems;

// Consecutive targets often share a bin (a grid of targets walks through a
// cell several targets at a time), so the lanes of a wavefront are grouped
// into runs of equal bins and the first lane of a run adds the run's length
// with one atomic.  heads: bit l set when lane l starts a run; lanes past T
// carry a bin no target has, so they start a run of their own.
struct BinRun {
    uint32_t bin;
    int lane, len;
    uint64_t heads;
    bool head;
};

__device__ __forceinline__ BinRun bin_run(const uint32_t *bins, int64_t t, int64_t T) {
    BinRun r;
    r.bin = t < T ? bins[t] : 0xffffffffu;
    r.lane = (int)(threadIdx.x & 63);
    const uint32_t prev = __shfl_up(r.bin, 1);   // lane 0 keeps its own value
    r.head = r.lane == 0 || prev != r.bin;
    r.heads = __ballot(r.head);
    const uint64_t after = r.lane == 63 ? 0 : r.heads >> (r.lane + 1);
    r.len = after ? __ffsll((long long)after) : 64 - r.lane;   // to the next head
    return r;
}

__global__ __launch_bounds__(256) void bin_hist_kernel(const uint32_t *bins, int64_t T,
                                                       int32_t *hist) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const BinRun r = bin_run(bins, t, T);
    if (t < T && r.head) atomicAdd(&hist[r.bin], r.len);
}

// Exclusive scan of a block's kScanThreads values in LDS (returned per
// thread) and their total.
__device__ __forceinline__ int32_t block_exclusive_scan(int32_t v, int32_t *lds, int32_t &total) {
    const int tid = threadIdx.x;
    lds[tid] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {   // Hillis-Steele inclusive scan
        const int32_t u = tid >= off ? lds[tid - off] : 0;
        __syncthreads();
        lds[tid] += u;
        __syncthreads();
    }
    total = lds[kScanThreads - 1];
    const int32_t r = lds[tid] - v;
    __syncthreads();
    return r;
}

// Pass 1: the sum of each chunk of kScanChunk counts.
__global__ __launch_bounds__(kScanThreads) void scan_sums_kernel(const int32_t *hist, int64_t nb,
                                                                  int32_t *sums) {
    __shared__ int32_t lds[kScanThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanChunk;
    int32_t v = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        const int64_t j = base + i * kScanThreads + threadIdx.x;   // coalesced
        v += j < nb ? hist[j] : 0;
    }
    int32_t total;
    (void)block_exclusive_scan(v, lds, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Pass 2 (one workgroup): exclusive scan of the chunk sums in place.
__global__ __launch_bounds__(kScanThreads) void scan_chunks_kernel(int32_t *sums, int64_t n) {
    __shared__ int32_t lds[kScanThreads];
    int32_t carry = 0;
    for (int64_t base = 0; base < n; base += kScanThreads) {
        const int64_t j = base + threadIdx.x;
        const int32_t v = j < n ? sums[j] : 0;
        int32_t total;
        const int32_t ex = block_exclusive_scan(v, lds, total);
        if (j < n) sums[j] = carry + ex;
        carry += total;
    }
}

// Pass 3: exclusive prefix of every count, from its chunk's offset.
__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(int32_t *hist, int64_t nb,
                                                                   const int32_t *sums) {
    __shared__ int32_t chunk[kScanChunk];
    __shared__ int32_t lds[kScanThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanChunk;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        const int64_t j = base + i * kScanThreads + threadIdx.x;
        chunk[i * kScanThreads + threadIdx.x] = j < nb ? hist[j] : 0;
    }
    __syncthreads();
    int32_t v[kScanItems];
    int32_t sum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {   // thread's own consecutive run
        v[i] = chunk[threadIdx.x * kScanItems + i];
        sum += v[i];
    }
    int32_t total;
    int32_t run = sums[blockIdx.x] + block_exclusive_scan(sum, lds, total);
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        chunk[threadIdx.x * kScanItems + i] = run;
        run += v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        const int64_t j = base + i * kScanThreads + threadIdx.x;
        if (j < nb) hist[j] = chunk[i * kScanThreads + threadIdx.x];
    }
}

// One atomic per run; the run's lanes take consecutive slots from the
// base its first lane received.
__global__ __launch_bounds__(256) void bin_scatter_kernel(const uint32_t *bins, int64_t T,
                                                          int32_t *offs, int32_t *perm) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const BinRun r = bin_run(bins, t, T);
    int32_t base = 0;
    if (t < T && r.head) base = atomicAdd(&offs[r.bin], r.len);
    // first lane of this lane's run: the highest head at or below it
    const uint64_t upto = r.lane == 63 ? r.heads : r.heads & ((2ull << r.lane) - 1);
    const int start = 63 - __clzll((long long)upto);
    base = __shfl(base, start);
    if (t < T) perm[base + (r.lane - start)] = (int32_t)t;
}

// Order-preserving selection of the targets of `perm` whose kd-tree
// replay is not pending (mark[t] == 0): keep flags, their exclusive scan
// (the same three passes), scatter.  Kernels only, like the order above.
__global__ void keep_kernel(const int32_t *perm, const uint8_t *mark, int64_t T, int32_t *pos) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < T) pos[i] = mark[perm[i]] == 0 ? 1 : 0;
}

__global__ void compact_kernel(const int32_t *perm, const uint8_t *mark, int64_t T,
                               const int32_t *pos, int32_t *out, int32_t *out_len) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T) return;
    const bool keep = mark[perm[i]] == 0;
    if (keep) out[pos[i]] = perm[i];
    if (i == T - 1) *out_len = pos[i] + (keep ? 1 : 0);
}

size_t select_workspace(int64_t T) {
    const int64_t chunks = (T + kScanChunk - 1) / kScanChunk;
    return sizeof(int32_t) * ((T + 63) / 64 * 64 + chunks);
}

hipError_t launch_select_unmarked(const int32_t *perm, const uint8_t *mark, int64_t T,
                                  int32_t *out, int32_t *out_len, void *tmp, size_t tmp_bytes,
                                  hipStream_t s) {
    if (T == 0) return hipSuccess;
    if (tmp_bytes < select_workspace(T)) return hipErrorInvalidValue;
    const int64_t chunks = (T + kScanChunk - 1) / kScanChunk;
    int32_t *pos = static_cast<int32_t *>(tmp);
    int32_t *sums = pos + (T + 63) / 64 * 64;
    const dim3 grid((unsigned)((T + 255) / 256));
    hipLaunchKernelGGL(keep_kernel, grid, dim3(256), 0, s, perm, mark, T, pos);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(scan_sums_kernel, dim3((unsigned)chunks), dim3(kScanThreads), 0, s, pos, T,
                       sums);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(scan_chunks_kernel, dim3(1), dim3(kScanThreads), 0, s, sums, chunks);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(scan_apply_kernel, dim3((unsigned)chunks), dim3(kScanThreads), 0, s, pos, T,
                       sums);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(compact_kernel, grid, dim3(256), 0, s, perm, mark, T, pos, out, out_len);
    return hipGetLastError();
}

namespace {

// Bins of a call: cells, merged in runs of 2^shift consecutive ranks while
// the grid has more than max(T, 65536) cells (a histogram far larger than
// the call would cost more to clear and scan than it saves).
int order_shift(int64_t ncells, int64_t T) {
    int shift = 0;
    while ((ncells - 1) >> shift >= std::max<int64_t>(T, 65536)) shift++;
    return shift;
}

int64_t order_bins(int64_t ncells, int64_t T) {
    return ((ncells - 1) >> order_shift(ncells, T)) + 1;
}

}  // namespace

size_t tile_sort_workspace(int64_t T, int64_t ncells) {
    const int64_t nb = order_bins(std::max<int64_t>(ncells, 1), T);
    const int64_t chunks = (nb + kScanChunk - 1) / kScanChunk;
    return sizeof(int32_t) * ((nb + 63) / 64 * 64 + chunks);
}

hipError_t launch_tile_order(int D, const double *targets, int64_t T, const SearchXf &xf,
                             const CellGridView &G, const uint32_t *cell_rank, uint32_t *bins,
                             int32_t *perm, void *tmp, size_t tmp_bytes, hipStream_t s) {
    if (T == 0) return hipSuccess;
    if (D != 2 && D != 3) return hipErrorInvalidValue;
    const int64_t ncells = (int64_t)G.nc[0] * G.nc[1] * G.nc[2];
    if (tmp_bytes < tile_sort_workspace(T, ncells)) return hipErrorInvalidValue;
    const int bshift = order_shift(ncells, T);
    const int64_t nb = ((ncells - 1) >> bshift) + 1;
    const int64_t chunks = (nb + kScanChunk - 1) / kScanChunk;
    int32_t *hist = static_cast<int32_t *>(tmp);
    int32_t *sums = hist + (nb + 63) / 64 * 64;
    const dim3 grid((unsigned)((T + 255) / 256));
    if (D == 2)
        hipLaunchKernelGGL(target_bin_kernel<2>, grid, dim3(256), 0, s, targets, T, xf, G,
                           cell_rank, bshift, bins);
    else
        hipLaunchKernelGGL(target_bin_kernel<3>, grid, dim3(256), 0, s, targets, T, xf, G,
                           cell_rank, bshift, bins);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if ((e = launch_fill_i32(hist, nb, 0, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(bin_hist_kernel, grid, dim3(256), 0, s, bins, T, hist);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(scan_sums_kernel, dim3((unsigned)chunks), dim3(kScanThreads), 0, s, hist, nb,
                       sums);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(scan_chunks_kernel, dim3(1), dim3(kScanThreads), 0, s, sums, chunks);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(scan_apply_kernel, dim3((unsigned)chunks), dim3(kScanThreads), 0, s, hist, nb,
                       sums);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(bin_scatter_kernel, grid, dim3(256), 0, s, bins, T, hist, perm);
    return hipGetLastError();
}

static size_t tile_lds_bytes(int D, int k, int per_octant) {
    const int slots = per_octant > 0 ? (per_octant << D) : k;
    return (size_t)slots * kTile * (sizeof(float) + sizeof(uint32_t)) + sizeof(RowTable);
}

hipError_t launch_tile_knn(int D, const TileParams &p, hipStream_t s) {
    if (p.T == 0) return hipSuccess;
    const bool oct = p.per_octant > 0;
    const size_t lds = tile_lds_bytes(D, p.k, p.per_octant);
    const dim3 grid((unsigned)((p.T + kTile - 1) / kTile));
    if (D == 2) {
        if (oct) hipLaunchKernelGGL((tile_knn_kernel<2, true>), grid, dim3(kTile), lds, s, p);
        else hipLaunchKernelGGL((tile_knn_kernel<2, false>), grid, dim3(kTile), lds, s, p);
    } else if (D == 3) {
        if (oct) hipLaunchKernelGGL((tile_knn_kernel<3, true>), grid, dim3(kTile), lds, s, p);
        else hipLaunchKernelGGL((tile_knn_kernel<3, false>), grid, dim3(kTile), lds, s, p);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace gs
