// Maxmin (farthest-point) ordering on the GPU (vecchia.rs:144-215): the
// reference's order, built in rounds that each append a batch of points.
//
// The reference appends, one point at a time, the point whose squared
// distance to the already ordered set ("key") is largest, ties to the larger
// index.  Keys only decrease, so after some points are ordered, the next ones
// follow from the candidates with the largest keys alone: sorted by
// (key, index) descending, candidate j is the next point exactly when no
// candidate ranked before it is closer to it than its key (its key is then
// unchanged after they are appended, and every other key is at most its own).
// A round therefore
//   1. updates the keys of the points around the previous batch (each point
//      of a batch can only lower keys within the square root of its own key),
//   2. collects the candidates at or above a threshold (key, index) and a
//      histogram of the keys below it, in 1/256-octave bins,
//   3. in one workgroup sorts the candidates and consumes them in
//      sub-rounds: append the longest prefix in which no candidate is closer
//      to an earlier one than its key, lower the other candidates' keys by
//      the appended points, drop those that fell below the threshold (every
//      point outside the candidate set is below it), re-sort the rest;
//      then it picks the next threshold from the histogram so that about the
//      candidate capacity is collected next time (never more by the
//      histogram's count, which bounds the true count from above).
// When more candidates than the capacity arrive anyway (ties of one key
// value, as on lattices), the round appends nothing and raises the threshold
// to the median (key, index) of the candidates it kept, which removes at
// least half of them.  Every step is exact (keys are sums of squares rounded
// as the reference's, compared as bit patterns), so the order equals the
// reference's.  The host drives the rounds in chunks and reads back the
// state between chunks; it finishes the traversal itself when batches stay
// small (tie-heavy inputs), from the keys the rounds left.
//
// Keys are stored per bucket slot as bits(key) + 1 (a non-negative double's
// bit pattern orders as the double), 0 once the point is ordered, so a
// 64-bit atomic minimum lowers a key and leaves ordered points alone.
// Compiled without contraction so squared distances round as the
// reference's.
#include "gs_internal.hpp"
#include "gs_maxmin.hpp"

namespace gs {

namespace {

constexpr int kBins = kMaxminBins;
constexpr int kCap = kMaxminCap;
constexpr int kBatchThreads = 1024;
constexpr int kBinShift = 44;      // bins of 1/256 octave: exponent + 8 mantissa bits
constexpr int kHash = 2048;        // spatial hash buckets of the batch kernel
constexpr uint16_t kNoId = 0xFFFF;
static_assert(kCap == 2 * kBatchThreads, "two candidates per thread in the batch kernel");
static_assert(kBins == kBatchThreads, "one histogram bin per thread in the batch kernel");
static_assert(kHash % kBatchThreads == 0, "whole hash buckets per thread");
static_assert(kCap < kNoId, "candidate ids fit 16 bits");

__device__ __forceinline__ unsigned long long key_bits(double d) {
    return (unsigned long long)__double_as_longlong(d);
}
__device__ __forceinline__ double bits_key(unsigned long long u) {
    return __longlong_as_double((long long)u);
}

template <int D>
__device__ __forceinline__ double sqd(const double *a, const double *b) {   // vecchia.rs:93-100
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

// Histogram bin of a key below the threshold key tk: bins of 1/256 octave
// counted down from tk's bin (the top 20 bits of the pattern: exponent and
// 8 mantissa bits), 4 octaves in all.  kb <= tk here.
__device__ __forceinline__ void hist_add(int *lh, unsigned long long tk, unsigned long long kb) {
    const unsigned long long k = (tk >> kBinShift) - (kb >> kBinShift);
    if (k < (unsigned long long)kBins) atomicAdd(&lh[k], 1);
}

__device__ __forceinline__ void flush_hist(int *lh, int *gh) {
    __syncthreads();
    for (int i = threadIdx.x; i < kBins; i += blockDim.x)
        if (lh[i]) atomicAdd(&gh[i], lh[i]);
}

// Keys to the first point, and the histogram of all of them below the
// initial threshold (a power of two above every squared distance).
template <int D>
__global__ __launch_bounds__(256) void maxmin_init_kernel(MaxminParams P) {
    __shared__ int lh[kBins];
    for (int i = threadIdx.x; i < kBins; i += blockDim.x) lh[i] = 0;
    __syncthreads();
    const unsigned long long tk = P.st->tau_key;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < P.n; q += stride) {
        if (q == (int64_t)P.first_slot) {
            P.enc[q] = 0;
            P.order[0] = (int64_t)P.id[q];
            continue;
        }
        const unsigned long long kb = key_bits(sqd<D>(P.cc + q * D, P.first));
        P.enc[q] = kb + 1;
        hist_add(lh, tk, kb);
    }
    flush_hist(lh, P.hist);
}

// Step 2: candidates at or above the threshold, histogram below it.
__global__ __launch_bounds__(256) void maxmin_collect_kernel(MaxminParams P) {
    __shared__ int lh[kBins];
    if (P.st->done) return;
    for (int i = threadIdx.x; i < kBins; i += blockDim.x) lh[i] = 0;
    __syncthreads();
    const unsigned long long tk = P.st->tau_key;
    const uint32_t ti = P.st->tau_idx;
    const int lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // the loop bound is uniform over the wavefront (ballot below)
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < P.n; base += stride) {
        const int64_t q = base + threadIdx.x;
        bool take = false;
        unsigned long long e = 0;
        if (q < P.n) {
            e = P.enc[q];
            if (e) {
                const unsigned long long kb = e - 1;
                if (kb > tk || (kb == tk && P.id[q] >= ti)) take = true;
                else hist_add(lh, tk, kb);
            }
        }
        const unsigned long long m = __ballot(take);
        if (m) {
            const int leader = __ffsll((long long)m) - 1;
            uint32_t pos0 = 0;
            if (lane == leader) pos0 = atomicAdd(&P.st->count, (uint32_t)__popcll(m));
            pos0 = __shfl(pos0, leader);
            if (take) {
                const uint32_t pos = pos0 + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                if (pos < (uint32_t)P.cap) {
                    P.cand_enc[pos] = e;
                    P.cand_slot[pos] = (uint32_t)q;
                }
            }
        }
    }
    flush_hist(lh, P.hist);
}

// Candidate id a ranks above id b: (key, index) descending; the id
// kNoId (padding) ranks below everything.
__device__ __forceinline__ bool ranks_above(const unsigned long long *ck, const uint32_t *ci,
                                            uint16_t a, uint16_t b) {
    if (a == kNoId) return false;
    if (b == kNoId) return true;
    return ck[a] > ck[b] || (ck[a] == ck[b] && ci[a] > ci[b]);
}

// Descending bitonic sort of candidate ids by (key, index), n2 a power of two.
__device__ void sort_ids(uint16_t *ord, const unsigned long long *ck, const uint32_t *ci, int n2) {
    for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = threadIdx.x; t < n2 / 2; t += blockDim.x) {
                const int i = 2 * t - (t & (j - 1));
                const int l = i + j;
                const uint16_t a = ord[i], b = ord[l];
                if ((i & k) == 0 ? ranks_above(ck, ci, b, a) : ranks_above(ck, ci, a, b)) {
                    ord[i] = b;
                    ord[l] = a;
                }
            }
            __syncthreads();
        }
    }
}

// Exclusive scan of one value per thread over the workgroup (wavefront
// shuffles, then the wavefront totals); *total gets the sum.  ws holds
// kBatchThreads / 64 values of LDS.
template <typename T>
__device__ T block_excl_scan(T v, T *ws, T *total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const T y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) ws[w] = x;
    __syncthreads();
    if (w == 0) {
        constexpr int nw = kBatchThreads / 64;
        T t = lane < nw ? ws[lane] : 0;
#pragma unroll
        for (int o = 1; o < nw; o <<= 1) {
            const T y = __shfl_up(t, o);
            if (lane >= o) t += y;
        }
        if (lane < nw) ws[lane] = t;
    }
    __syncthreads();
    const T r = x - v + (w > 0 ? ws[w - 1] : 0);
    *total = ws[kBatchThreads / 64 - 1];
    __syncthreads();
    return r;
}

__device__ __forceinline__ bool el_gt(unsigned long long ka, uint32_t ia, unsigned long long kb, uint32_t ib) {
    return ka > kb || (ka == kb && ia > ib);
}

// Descending bitonic sort of 2 * kBatchThreads (key, index, id) elements
// held two per thread (positions 2 tid, 2 tid + 1): stages whose partner is
// in the same wavefront exchange through shuffles, the others through LDS
// (xk, xi, xid: 2 * kBatchThreads entries each).  Equal elements (padding)
// may be exchanged freely.
__device__ void sort_regs(unsigned long long (&k)[2], uint32_t (&i)[2], uint32_t (&id)[2],
                          unsigned long long *xk, uint32_t *xi, uint32_t *xid) {
    const int tid = threadIdx.x;
    constexpr int N = 2 * kBatchThreads;
    for (int kk = 2; kk <= N; kk <<= 1) {
        for (int j = kk >> 1; j > 0; j >>= 1) {
            if (j == 1) {
                const bool desc = ((2 * tid) & kk) == 0;
                if (desc == el_gt(k[1], i[1], k[0], i[0])) {
                    const unsigned long long tk = k[0];
                    k[0] = k[1];
                    k[1] = tk;
                    const uint32_t t1 = i[0];
                    i[0] = i[1];
                    i[1] = t1;
                    const uint32_t t2 = id[0];
                    id[0] = id[1];
                    id[1] = t2;
                }
                continue;
            }
            unsigned long long ok[2];
            uint32_t oi[2], oid[2];
            if (j < 128) {
                const int m = j >> 1;
#pragma unroll
                for (int b = 0; b < 2; b++) {
                    ok[b] = __shfl_xor(k[b], m);
                    oi[b] = __shfl_xor(i[b], m);
                    oid[b] = __shfl_xor(id[b], m);
                }
            } else {
#pragma unroll
                for (int b = 0; b < 2; b++) {
                    xk[2 * tid + b] = k[b];
                    xi[2 * tid + b] = i[b];
                    xid[2 * tid + b] = id[b];
                }
                __syncthreads();
#pragma unroll
                for (int b = 0; b < 2; b++) {
                    const int q = (2 * tid + b) ^ j;
                    ok[b] = xk[q];
                    oi[b] = xi[q];
                    oid[b] = xid[q];
                }
                __syncthreads();
            }
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const int pos = 2 * tid + b;
                const bool keep_greater = ((pos & j) == 0) == ((pos & kk) == 0);
                const bool take = keep_greater ? el_gt(ok[b], oi[b], k[b], i[b]) : el_gt(k[b], i[b], ok[b], oi[b]);
                if (take) {
                    k[b] = ok[b];
                    i[b] = oi[b];
                    id[b] = oid[b];
                }
            }
        }
    }
}

// Spatial hash of the candidates: cells of side h, more than the square root
// of every candidate key, so a candidate that can lower another's key lies in
// one of its 3^D neighbour cells; cells wrap onto a grid of kHash buckets
// (64 x 32 in 2-D, 16 x 16 x 8 in 3-D), so a bucket may hold several far-apart
// cells, which only adds distance tests.
template <int D>
__device__ __forceinline__ int wrap_bucket(int cx, int cy, int cz) {
    if (D == 2) return (cx & 63) | ((cy & 31) << 6);
    return (cx & 15) | ((cy & 15) << 4) | ((cz & 7) << 8);
}

// Step 3 (one workgroup): the candidates, sorted, are consumed in
// sub-rounds.  Each sub-round appends the longest prefix of the pending
// candidates in which none is closer to an earlier one than its key, lowers
// the keys of the other pending candidates by the appended points, drops
// those that fell below the threshold (every point outside the candidate set
// is below it, so a pending candidate at or above it is still ahead of all
// of them), and merges the lowered ones back into rank order.  Then the
// round's batch is committed and the next threshold is picked.
template <int D>
__global__ __launch_bounds__(kBatchThreads) void maxmin_batch_kernel(MaxminParams P) {
    static_assert(kHash == 2048, "wrapped grid of 2048 buckets");
    __shared__ unsigned long long ck[kCap];   // per candidate id: bits(key) + 1
    __shared__ uint32_t ci[kCap], cs[kCap];   // original index, slot
    __shared__ double cx[kCap * D];           // coordinates
    __shared__ int cc3[kCap * D];             // hash cell coordinates
    __shared__ uint16_t ord[kCap];            // pending ids in rank order
    __shared__ int16_t rnk[kCap];             // rank of a pending id
    __shared__ int16_t stamp[kCap];           // -1 pending, -2 dropped, t: appended in sub-round t
    __shared__ uint16_t bsel[kCap];           // appended ids in order
    __shared__ uint16_t tmpu[kCap], tmpg[kCap];
    __shared__ int hcnt[kHash];               // bucket ends, then starts
    __shared__ int lhist[kBins];
    __shared__ int wsum[kBatchThreads / 64];
    __shared__ int64_t wsum64[kBatchThreads / 64];
    __shared__ uint8_t chg[kCap];              // pending key lowered this sub-round
    __shared__ int sh_b, sh_k1, sh_k2, sh_kf;
    MaxminState *st = P.st;
    const int tid = threadIdx.x;
    if (st->done) return;
    const uint32_t C = st->count;
    const unsigned long long tk = st->tau_key;
    const uint32_t ti = st->tau_idx;
    const int64_t s0 = st->s;
    const int cap = P.cap;
    const int cl = C < (uint32_t)cap ? (int)C : cap;
    {
        // candidates in registers, two per thread, sorted; the id of a
        // candidate is its collection position until the hash reorders them
        unsigned long long rk[2];
        uint32_t ri[2], rid[2];
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int j = 2 * tid + e;
            if (j < cl) {
                const uint32_t slot = P.cand_slot[j];
                rk[e] = P.cand_enc[j];
                ri[e] = P.id[slot];
                rid[e] = (uint32_t)j;
                ck[j] = rk[e];
                cs[j] = slot;
                ci[j] = ri[e];
            } else {
                rk[e] = 0;   // below every candidate (their enc is >= 1)
                ri[e] = 0;
                rid[e] = kNoId;
            }
        }
        for (int i = tid; i < kBins; i += kBatchThreads) lhist[i] = P.hist[i];
        if (cl > 1) {
            unsigned long long *xk = reinterpret_cast<unsigned long long *>(cx);
            uint32_t *xi = reinterpret_cast<uint32_t *>(xk + kCap);
            uint32_t *xid = xi + kCap;
            sort_regs(rk, ri, rid, xk, xi, xid);
        }
#pragma unroll
        for (int e = 0; e < 2; e++) ord[2 * tid + e] = (uint16_t)rid[e];
        __syncthreads();
    }
    if (C > (uint32_t)cap) {
        // too many candidates: nothing is appended; the median of the kept
        // ones becomes the threshold (strictly above the old one, since the
        // kept candidates are distinct and at or above it)
        for (int i = tid; i < kBins; i += kBatchThreads) P.hist[i] = 0;
        if (tid == 0) {
            const uint16_t m = ord[cap / 2];
            st->tau_key = ck[m] - 1;
            st->tau_idx = ci[m];
            st->count = 0;
            st->batch = 0;
            st->rows = 0;
            st->rounds += 1;
            st->retries += 1;
        }
        return;
    }
    // hash over every candidate; cells sized by the largest key, widened so
    // that a coordinate difference below sqrt(key) spans at most one cell
    // boundary after the rounded scaling, and at least 2^-30 of the extent
    // (then the scaled coordinates stay below 2^30 and their rounding error
    // below 2^-22 of a cell)
    double ext = 0.0;
#pragma unroll
    for (int d = 0; d < D; d++) ext = fmax(ext, (double)P.nn[d] * P.cell);
    const double h = fmax(cl > 0 ? sqrt(bits_key(ck[ord[0]] - 1)) * (1.0 + 1e-6) : 1.0,
                          ext * 0x1p-30) + 1e-300;
    const double inv_h = 1.0 / h;
    for (int b = tid; b < kHash; b += kBatchThreads) hcnt[b] = 0;
    for (int j = tid; j < cl; j += kBatchThreads) {
        const int64_t slot = cs[j];
#pragma unroll
        for (int d = 0; d < D; d++) {
            const double x = P.cc[slot * D + d];
            cx[j * D + d] = x;
            const int64_t c = sat_i64((x - P.origin[d]) * inv_h);
            cc3[j * D + d] = (int)(c < 0 ? 0 : (c > 0x7FFFFFF0 ? 0x7FFFFFF0 : c));
        }
    }
    __syncthreads();
    for (int j = tid; j < cl; j += kBatchThreads)
        atomicAdd(&hcnt[wrap_bucket<D>(cc3[j * D], cc3[j * D + 1], D == 3 ? cc3[j * D + 2] : 0)], 1);
    __syncthreads();
    {
        constexpr int per = kHash / kBatchThreads;
        int v[per];
        int tot = 0;
#pragma unroll
        for (int e = 0; e < per; e++) { v[e] = hcnt[tid * per + e]; tot += v[e]; }
        int all;
        int run = block_excl_scan(tot, wsum, &all);
#pragma unroll
        for (int e = 0; e < per; e++) { run += v[e]; hcnt[tid * per + e] = run; }
    }
    __syncthreads();
    // candidates move to bucket order: the id of a candidate is from here on
    // its position in the hash, so a bucket is a contiguous id range
    {
        int pos[2] = {0, 0};
        unsigned long long k2[2];
        uint32_t i2[2], s2[2];
        double x2[2][D];
        int c2[2][D];
        for (int e = 0; e < 2; e++) {
            const int j = tid * 2 + e;
            if (j >= cl) break;
            const int b = wrap_bucket<D>(cc3[j * D], cc3[j * D + 1], D == 3 ? cc3[j * D + 2] : 0);
            pos[e] = atomicSub(&hcnt[b], 1) - 1;
            k2[e] = ck[j];
            i2[e] = ci[j];
            s2[e] = cs[j];
#pragma unroll
            for (int d = 0; d < D; d++) { x2[e][d] = cx[j * D + d]; c2[e][d] = cc3[j * D + d]; }
        }
        __syncthreads();
        for (int e = 0; e < 2; e++) {
            const int j = tid * 2 + e;
            if (j >= cl) break;
            const int q = pos[e];
            ck[q] = k2[e];
            ci[q] = i2[e];
            cs[q] = s2[e];
#pragma unroll
            for (int d = 0; d < D; d++) { cx[q * D + d] = x2[e][d]; cc3[q * D + d] = c2[e][d]; }
            tmpu[j] = (uint16_t)q;     // old id -> new id
        }
        __syncthreads();
        for (int p = tid; p < cl; p += kBatchThreads) {
            const uint16_t q = tmpu[ord[p]];
            ord[p] = q;
            rnk[q] = (int16_t)p;
            stamp[q] = -1;
            chg[q] = 0;
        }
        __syncthreads();
    }
    // visits the ids hashed near candidate j (its 3^D neighbour cells)
    auto near = [&](int j, auto &&fn) {
        const int c0 = cc3[j * D], c1 = cc3[j * D + 1], c2 = D == 3 ? cc3[j * D + 2] : 0;
        for (int dz = (D == 3 ? -1 : 0); dz <= (D == 3 ? 1 : 0); dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    const int b = wrap_bucket<D>(c0 + dx, c1 + dy, c2 + dz);
                    const int e = b + 1 < kHash ? hcnt[b + 1] : cl;
                    for (int q = hcnt[b]; q < e; q++) fn(q);
                }
    };
    // dense when the candidates crowd few cells (large keys against the
    // spread of the candidates): the neighbour lists are then most of the
    // candidates, and direct scans over the ranks are cheaper
    bool dense;
    {
        int nearcnt = 0;
        for (int e = 0; e < 2; e++) {
            const int j = 2 * tid + e;
            if (j >= cl) break;
            near(j, [&](int) { nearcnt++; });
        }
        int all;
        (void)block_excl_scan(nearcnt, wsum, &all);
        dense = all > 48 * cl;
    }
    int np = cl, nb = 0, nsub = 0;
    for (int t = 0; np > 0; t++) {
        nsub = t + 1;
        if (tid == 0) sh_b = np;
        __syncthreads();
        // first rank with a closer pending candidate ranked before it, a
        // chunk of ranks at a time (dense: 64 ranks, 16 threads per rank
        // scanning the ranks before it)
        for (int p0 = 1; dense && p0 < np; p0 += kBatchThreads / 16) {
            const int p = p0 + (tid >> 4);
            if (p < np) {
                const int j = ord[p];
                const double kj = bits_key(ck[j] - 1);
                for (int r = tid & 15; r < p; r += 16) {
                    if (sqd<D>(cx + j * D, cx + ord[r] * D) < kj) {
                        atomicMin(&sh_b, p);
                        break;
                    }
                }
            }
            __syncthreads();
            const int bnow = sh_b;
            __syncthreads();
            if (bnow < np) break;
        }
        for (int p0 = 1; !dense && p0 < np; p0 += kBatchThreads) {
            const int p = p0 + tid;
            if (p < np) {
                const int j = ord[p];
                const double kj = bits_key(ck[j] - 1);
                bool conf = false;
                near(j, [&](int i) {
                    if (!conf && stamp[i] == -1 && rnk[i] < p && sqd<D>(cx + j * D, cx + i * D) < kj) conf = true;
                });
                if (conf) atomicMin(&sh_b, p);
            }
            __syncthreads();
            const int bnow = sh_b;
            __syncthreads();
            if (bnow < np) break;
        }
        const int bb = sh_b;
        for (int p = tid; p < bb; p += kBatchThreads) {
            const int j = ord[p];
            stamp[j] = (int16_t)t;
            bsel[nb + p] = (uint16_t)j;
        }
        nb += bb;
        __syncthreads();
        // the appended points lower the keys of the pending candidates near
        // them: pushed through the hash, or (few appended points, or dense)
        // pulled by each pending candidate from the appended list below
        const bool pull = dense || bb <= 8;
        for (int p = tid; !pull && p < bb; p += kBatchThreads) {
            const int a = ord[p];
            near(a, [&](int q) {
                if (stamp[q] == -1) {
                    const unsigned long long ne = key_bits(sqd<D>(cx + q * D, cx + a * D)) + 1;
                    if (ne < ck[q]) {
                        atomicMin(&ck[q], ne);
                        chg[q] = 1;
                    }
                }
            });
        }
        __syncthreads();
        // the other pending candidates: unchanged (still in rank order),
        // lowered, or dropped (below the threshold); positions bb + 2 tid,
        // bb + 2 tid + 1 per thread
        uint16_t keep_u[2], keep_g[2];
        int nu = 0, ng = 0;
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int p = bb + 2 * tid + e;
            if (p >= np) break;
            const int j = ord[p];
            if (pull) {
                unsigned long long en = ck[j];
                for (int r = 0; r < bb; r++) {
                    const unsigned long long ne = key_bits(sqd<D>(cx + j * D, cx + ord[r] * D)) + 1;
                    en = ne < en ? ne : en;
                }
                if (en < ck[j]) {
                    ck[j] = en;
                    chg[j] = 1;
                }
            }
            if (!chg[j]) {
                keep_u[nu++] = (uint16_t)j;
                continue;
            }
            chg[j] = 0;
            const unsigned long long kb = ck[j] - 1;
            if (!(kb > tk || (kb == tk && ci[j] >= ti))) stamp[j] = -2;
            else keep_g[ng++] = (uint16_t)j;
        }
        int mu, mg;
        const int pu = block_excl_scan(nu, wsum, &mu);
        const int pg = block_excl_scan(ng, wsum, &mg);
        for (int e = 0; e < nu; e++) tmpu[pu + e] = keep_u[e];
        for (int e = 0; e < ng; e++) tmpg[pg + e] = keep_g[e];
        np = mu + mg;
        if (np == 0) break;
        int g2 = 1;
        while (g2 < mg) g2 <<= 1;
        for (int q = mg + tid; q < g2; q += kBatchThreads) tmpg[q] = kNoId;
        __syncthreads();
        if (mg > 1) sort_ids(tmpg, ck, ci, g2);
        // merge: an element's rank is its index plus the number of elements
        // of the other list that rank above it
        for (int a = tid; a < mu; a += kBatchThreads) {
            const uint16_t id = tmpu[a];
            int lo = 0, hi = mg;    // first g that does not rank above id
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (ranks_above(ck, ci, tmpg[mid], id)) lo = mid + 1;
                else hi = mid;
            }
            ord[a + lo] = id;
            rnk[id] = (int16_t)(a + lo);
        }
        for (int g = tid; g < mg; g += kBatchThreads) {
            const uint16_t id = tmpg[g];
            int lo = 0, hi = mu;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (ranks_above(ck, ci, tmpu[mid], id)) lo = mid + 1;
                else hi = mid;
            }
            ord[g + lo] = id;
            rnk[id] = (int16_t)(g + lo);
        }
        __syncthreads();
    }
    const int b = nb;
    const bool last = s0 + b >= P.n;
    // dropped candidates go back into the histogram with their new keys
    for (int j = tid; j < cl; j += kBatchThreads)
        if (stamp[j] == -2) hist_add(lhist, tk, ck[j] - 1);
    // commit the batch and the update boxes of its points: every point
    // whose key the batch can lower lies within sqrt(key) of a batch point
    // (every remaining key is at most the batch's smallest); the box is
    // widened for rounding as the host traversal's
    int64_t rows2[2] = {0, 0};
    for (int e = 0; e < 2; e++) {
        const int k = 2 * tid + e;
        if (k >= b) break;
        const int j = bsel[k];
        P.order[s0 + k] = (int64_t)ci[j];
        P.enc[cs[j]] = 0;
        P.bslot[k] = cs[j];
        const double rad = sqrt(bits_key(ck[j] - 1)) * (1.0 + 1e-9) + 1e-300;
        int64_t rows = 1;
#pragma unroll
        for (int d = 0; d < D; d++) {
            int64_t lo = sat_i64((cx[j * D + d] - rad - P.origin[d]) / P.cell);
            int64_t hi = sat_i64((cx[j * D + d] + rad - P.origin[d]) / P.cell);
            lo = lo < 0 ? 0 : (lo > P.nn[d] - 1 ? P.nn[d] - 1 : lo);
            hi = hi < 0 ? 0 : (hi > P.nn[d] - 1 ? P.nn[d] - 1 : hi);
            P.bbox[k * 6 + d] = (int32_t)lo;
            P.bbox[k * 6 + 3 + d] = (int32_t)hi;
            if (d > 0) rows *= hi - lo + 1;
        }
        rows2[e] = rows;
    }
    // exclusive scan of the rows over the batch (two entries per thread)
    int64_t total_rows;
    {
        const int64_t excl = block_excl_scan<int64_t>(rows2[0] + rows2[1], wsum64, &total_rows);
        if (2 * tid < b) P.brow[2 * tid] = excl;
        if (2 * tid + 1 < b) P.brow[2 * tid + 1] = excl + rows2[0];
        if (tid == 0) P.brow[b] = total_rows;
    }
    // next threshold: histogram bins from the top down until the capacity
    // is covered (the counts bound the next round's candidates from above)
    const int target = cap;
    const int64_t hv = tid < kBins ? lhist[tid] : 0;
    if (tid == 0) {
        sh_k1 = kBins;
        sh_k2 = kBins;
        sh_kf = kBins;
    }
    int64_t hall;
    const int64_t cum = block_excl_scan<int64_t>(hv, wsum64, &hall) + hv;   // bins 0..tid
    if (tid < kBins) {
        if (cum >= target) atomicMin(&sh_k1, tid);
        if (cum > cap) atomicMin(&sh_k2, tid);
        if (cum > 0) atomicMin(&sh_kf, tid);
    }
    __syncthreads();
    for (int i = tid; i < kBins; i += kBatchThreads) P.hist[i] = 0;
    if (tid == 0) {
        int chosen = min(sh_k1 < kBins ? sh_k1 : kBins - 1, sh_k2 - 1);
        // the chosen bins are empty: take the first nonempty bin even when it
        // alone exceeds cap (the overflow step then splits it), or drop the
        // whole histogram range when every bin is empty
        if (chosen < sh_kf) chosen = sh_kf < kBins ? sh_kf : kBins - 1;
        const unsigned long long hb = tk >> kBinShift;
        st->tau_key = hb >= (unsigned long long)chosen ? (hb - (unsigned long long)chosen) << kBinShift : 0;
        st->tau_idx = 0;
        st->count = 0;
        st->batch = b;
        st->rows = last ? 0 : total_rows;
        st->s = s0 + b;
        st->done = last ? 1 : 0;
        st->rounds += 1;
        st->subrounds += nsub;
    }
}

// Step 1: lower the keys around the last batch.  One wavefront per row of
// buckets (a contiguous slot range) of some batch point's box.
template <int D>
__global__ __launch_bounds__(256) void maxmin_update_kernel(MaxminParams P) {
    const int64_t R = P.st->rows;
    const int b = P.st->batch;
    if (R <= 0 || b <= 0) return;
    const int lane = threadIdx.x & 63;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; r < R; r += nw) {
        int lo = 0, hi = b - 1;   // last batch point j with brow[j] <= r
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (P.brow[mid] <= r) lo = mid;
            else hi = mid - 1;
        }
        const int j = lo;
        const int64_t rr = r - P.brow[j];
        const int32_t *bx = P.bbox + j * 6;
        int64_t b0;
        if (D == 2) {
            b0 = ((int64_t)bx[1] + rr) * P.strides[1];
        } else {
            const int64_t ny = (int64_t)bx[4] - bx[1] + 1;
            b0 = ((int64_t)bx[1] + rr % ny) * P.strides[1] + ((int64_t)bx[2] + rr / ny) * P.strides[2];
        }
        const int64_t q0 = P.start[b0 + bx[0]], q1 = P.start[b0 + bx[3] + 1];
        const int64_t ps = P.bslot[j];
        double pp[D];
#pragma unroll
        for (int d = 0; d < D; d++) pp[d] = P.cc[ps * D + d];
        for (int64_t q = q0 + lane; q < q1; q += 64) {
            const unsigned long long e = P.enc[q];
            if (e == 0) continue;
            const unsigned long long ne = key_bits(sqd<D>(P.cc + q * D, pp)) + 1;
            if (ne < e) atomicMin(&P.enc[q], ne);
        }
    }
}

}  // namespace

hipError_t launch_maxmin_init(int D, const MaxminParams &p, hipStream_t s) {
    const int64_t want = (p.n + 255) / 256;
    const unsigned blocks = (unsigned)(want < 2048 ? (want > 0 ? want : 1) : 2048);
    if (D == 2) hipLaunchKernelGGL((maxmin_init_kernel<2>), dim3(blocks), dim3(256), 0, s, p);
    else if (D == 3) hipLaunchKernelGGL((maxmin_init_kernel<3>), dim3(blocks), dim3(256), 0, s, p);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_maxmin_round(int D, const MaxminParams &p, hipStream_t s) {
    if (D != 2 && D != 3) return hipErrorInvalidValue;
    if (p.cap < 2 || p.cap > kCap || (p.cap & (p.cap - 1)) != 0) return hipErrorInvalidValue;
    // a few points per thread: fewer blocks flush fewer histogram atomics
    const int64_t want = (p.n + 1023) / 1024;
    const unsigned cblocks = (unsigned)(want < 512 ? (want > 0 ? want : 1) : 512);
    if (D == 2) hipLaunchKernelGGL((maxmin_update_kernel<2>), dim3(1024), dim3(256), 0, s, p);
    else hipLaunchKernelGGL((maxmin_update_kernel<3>), dim3(1024), dim3(256), 0, s, p);
    hipLaunchKernelGGL(maxmin_collect_kernel, dim3(cblocks), dim3(256), 0, s, p);
    if (D == 2) hipLaunchKernelGGL((maxmin_batch_kernel<2>), dim3(1), dim3(kBatchThreads), 0, s, p);
    else hipLaunchKernelGGL((maxmin_batch_kernel<3>), dim3(1), dim3(kBatchThreads), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_maxmin_update(int D, const MaxminParams &p, hipStream_t s) {
    if (D == 2) hipLaunchKernelGGL((maxmin_update_kernel<2>), dim3(1024), dim3(256), 0, s, p);
    else if (D == 3) hipLaunchKernelGGL((maxmin_update_kernel<3>), dim3(1024), dim3(256), 0, s, p);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_maxmin_batch(int D, const MaxminParams &p, hipStream_t s) {
    if (D == 2) hipLaunchKernelGGL((maxmin_batch_kernel<2>), dim3(1), dim3(kBatchThreads), 0, s, p);
    else if (D == 3) hipLaunchKernelGGL((maxmin_batch_kernel<3>), dim3(1), dim3(kBatchThreads), 0, s, p);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace gs
