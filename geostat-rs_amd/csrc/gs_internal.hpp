// Shared declarations between the host runtime (gs_host.cpp) and the
// kernel launchers (gs_search.hip, gs_solve.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gs_device.hpp"

struct gs_vplan;
// Vecchia plan properties for the MLE driver (defined with the plan).
int gs_plan_dim(const gs_vplan *p);
int64_t gs_plan_points(const gs_vplan *p);
double gs_plan_extent(const gs_vplan *p);
// Records msg as the calling thread's last error and returns code.
int32_t gs_fail(int32_t code, const char *msg);

namespace gs {

// Search frame: search_point (kriging.rs:16-21) applies
// Anisotropy::transform_vector to absolute coordinates.  The 2-D form keeps
// the reference's exact expression [x s + y c, (x c - y s) / ratio]
// (model.rs:362-364) and the 3-D form the GSLIB matrix rows summed left to
// right (model.rs:358-360), so search distances are bit-identical.
struct SearchXf {
    int32_t on;
    int32_t pad;
    double s, c, ratio;
    double r[9];
};

// Uniform cell grid over the search-frame data coordinates (CSR by cell,
// x fastest), used by the tile search.
struct CellGridView {
    double origin[3];
    double cell;
    double inv_cell;
    int32_t nc[3];
    int32_t pad;
    const int32_t *start;    // [ncells + 1]
    const double *pts;       // [n][D] search-frame coords in cell order
    const uint32_t *idx;     // [n] original indices in cell order
};

struct TileParams {
    CellGridView G;
    const double *targets;   // [T][D] original frame
    int64_t T;
    const int32_t *perm;     // [T] targets in Hilbert order of their cell
    int32_t k;               // heap mode list length
    int32_t per_octant;      // > 0: octant mode
    int32_t max_nb;          // octant mode truncation
    int32_t stride;
    double r2;
    SearchXf xf;
    int32_t *out_idx;
    int32_t *out_count;
    int32_t *flag_list;      // targets for the kd-tree replay (ties, scattered tiles)
    int32_t *flag_len;
    uint8_t *flag_mark;      // optional [T]: set to 1 for every flagged target
};

struct KnnParams {
    const int32_t *list;     // optional target ids (exact replay of flagged targets)
    const int32_t *list_len; // device counter for list
    const double *tcoords;   // [n][D] search-frame coords in kd-tree layout order
    const uint32_t *tidx;    // kd-tree layout position -> original index
    int64_t n;
    const double *targets;   // [T][D] original frame
    int64_t T;
    int32_t k;               // heap mode: max_neighbors (or n)
    int32_t per_octant;      // > 0: octant mode
    int32_t max_nb;          // octant mode: truncation length
    int32_t depth;           // kd-tree depth bound (stack size)
    double r2;               // radius^2 or +inf
    SearchXf xf;
    int32_t *out_idx;        // [T][stride]
    int32_t *out_count;      // [T]
    int32_t stride;
    // Large neighbourhoods: heaps in device memory ([block][slot][lane]
    // instead of LDS; the grid must not exceed heap_blocks), and outputs
    // indexed by list position instead of target id.
    double *gheap_d;
    uint32_t *gheap_ix;
    int64_t heap_blocks;
    int32_t out_by_pos;
};

struct SolveParams {
    const double *coords;    // [n][D] original frame
    const double *values;    // [n]
    const double *targets;   // [T][D]
    const int32_t *nbr;      // [T][stride]
    const int32_t *count;    // [T]
    const int32_t *list;     // optional bucket list of target ids
    const int32_t *list_len; // device counter for list
    int64_t T;
    int32_t stride;
    int32_t min_nb;
    int32_t method;
    int32_t degree;
    double sk_mean;
    double center[3];
    double scale;
    const DevModel *model;   // device memory
    double *value;
    double *variance;
    double *lagrange;
    int32_t *status;
    // Extended systems (kernels instantiated with X = true only):
    const double *err;       // [n] measurement-error variances on the data diagonal, or null
    const double *xdrift;    // [n][M-1] normalized external covariates at the data, or null
    const double *tdrift;    // [T][M-1] covariates at the targets (raw)
    double xmean[3], xscale[3];
    const double *offsets;   // [n_off][D] block discretization, or null (point kriging)
    int32_t n_off;
    const double *cbb;       // device scalar: within-block covariance
    // FP32-refined solves: systems whose FP32 factorization is not trusted
    // (smallest pivot below fb_ratio * N * matrix scale, N the system size)
    // are appended to fb_list (count in *fb_len) for the FP64 kernel.
    int32_t *fb_list;
    int32_t *fb_len;
    double fb_ratio;
};

// Global neighbourhood above 64 points: one shared LU, per-target
// substitutions (gs_global.hip).
struct GlobalParams {
    const double *coords;    // [n][D]
    const double *values;    // [n]
    int64_t n, N;            // data points, system size n + M
    int32_t M, method;       // drift terms; 0 simple, 1 ordinary, 2 universal, 3 external
    double sk_mean;
    double center[3];
    double scale;
    const double *err;       // [n] measurement-error variances or null
    const double *xdrift;    // [n][nv] normalized covariates or null
    int32_t nv;
    double xmean[3], xscale[3];
    const double *tdrift;    // [T][nv] target covariates
    const DevModel *model;
    double *lu;              // [N][N] factors (row-major, L below the diagonal)
    int32_t *piv;            // [N] row swapped with row k at step k
    int32_t *singular;       // device flag: the factorization hit a tiny pivot
    const double *targets;   // [T][D]
    int64_t T, t0, Tc;       // all targets, first of this chunk, chunk size
    int64_t stride;          // scratch row stride (>= Tc)
    double *w, *b0;          // scratch [N][stride]
    double *value, *variance, *lagrange;
    int32_t *status;
    const double *offsets;   // block discretization or null
    int32_t n_off;
    const double *cbb;
};

hipError_t launch_global_factor(int D, bool lean, GlobalParams p, double *d_scale, hipStream_t s);
hipError_t launch_global_solve(int D, bool lean, const GlobalParams &p, hipStream_t s);

// Nearest predecessors of the Vecchia plan (gs_plan.hip): canonical m
// nearest among positions < k for every k, over the reference's bucket grid.
struct PredParams {
    const double *pts;        // [n][D] coordinates in ordered position
    int64_t n;
    int32_t m;
    int64_t brute;            // positions k <= brute scan all predecessors
    double origin[3];
    double cell;
    int64_t nn[3], strides[3];
    const int64_t *start;     // [buckets + 1] CSR bucket offsets
    const int32_t *items;     // positions, ascending inside each bucket
    int32_t *out;             // [n][m] positions, ascending (d2, position), -1 padded
    int32_t *count;           // [n]
    int32_t *flag;            // [n] 1: a tie at the m-th distance (replay on the host)
};

hipError_t launch_pred_knn(int D, const PredParams &p, hipStream_t s);

struct VecchiaParams {
    const double *coords;    // [n][D]
    const double *zc;        // [n] centred values
    const int64_t *order;    // [n]
    const int32_t *nb;       // [n][m] rows in ordered position
    const int32_t *nb_count; // [n]
    int64_t n;
    int32_t m;
    int32_t n_models;
    const DevModel *models;  // [n_models] in device memory
    double *partial;         // [n_blocks][n_models]
    int32_t *status;         // [n_models]
    int32_t points_per_block;
    // FP32-refined batches: refine_fail[p] = 1 when a conditional of model p
    // failed the FP32 pivot test (the model is recomputed in FP64)
    int32_t *refine_fail;
    double fb_ratio;
};

hipError_t launch_knn(int D, const KnnParams &p, hipStream_t s);
// In-radius point count of each listed target (kd-tree traversal, no heap),
// written by list position.
hipError_t launch_radius_count(int D, const KnnParams &p, int32_t *counts, hipStream_t s);

// Moving neighbourhoods above 64 points (gs_big.hip): one workgroup per
// listed target builds the bordered system from its neighbour list and
// solves it by LU with partial pivoting in the reference's order.
struct BigParams {
    SolveParams S;           // data, model, method, extended inputs, outputs (by target id)
    const int32_t *list;     // [F] target ids
    int64_t F;
    const int32_t *nbr;      // [F][stride] neighbour lists by list position
    const int32_t *count;    // [F]
    int32_t stride;
    int32_t M;               // drift terms
    int32_t nv;              // external covariates
    int32_t cap;             // largest system the scratch holds (neighbours)
    double *scratch;         // [blocks][slot]
    int64_t slot;            // doubles per block: (cap+M)^2 + 2 (cap+M)
};
hipError_t launch_big_solve(int D, bool lean, const BigParams &p, int64_t blocks, hipStream_t s);
// Targets whose neighbour count exceeds `limit`, appended to list (length
// in *len).
hipError_t launch_collect_over(const int32_t *count, int64_t T, int32_t limit, int32_t *list,
                               int32_t *len, hipStream_t s);
// *out = max of counts[0..n) (counts >= 0).
hipError_t launch_max_i32(const int32_t *counts, int64_t n, int32_t *out, hipStream_t s);
// Hilbert rank of every cell of the grid, [ncells] (build time, synchronous).
hipError_t build_cell_ranks(int D, const CellGridView &G, uint32_t *rank, hipStream_t s);
// Target processing order: perm = targets sorted by the Hilbert rank of
// their cell; bins is [T] scratch, tmp holds tile_sort_workspace bytes.
size_t tile_sort_workspace(int64_t T, int64_t ncells);
hipError_t launch_tile_order(int D, const double *targets, int64_t T, const SearchXf &xf,
                             const CellGridView &G, const uint32_t *cell_rank, uint32_t *bins,
                             int32_t *perm, void *tmp, size_t tmp_bytes, hipStream_t s);
hipError_t launch_tile_knn(int D, const TileParams &p, hipStream_t s);
hipError_t launch_fill_i32(int32_t *p, int64_t n, int32_t value, hipStream_t s);
// Returns hipErrorInvalidValue when (D, NB, M) has no instantiation.
// variant: group-size choice (0/1); lean: every structure is spherical,
// exponential, Gaussian or closed-form Matern.  Returns
// hipErrorInvalidDeviceFunction for a (variant, class) pair that is not
// compiled (general-class kernels exist for variant 0 only).
// ext: the extended-system kernels (measurement error, external drift,
// block kriging; variant 0 only).
hipError_t launch_krige_fp64(int D, int NB, int M, int variant, bool lean, bool ext,
                             const SolveParams &p, int64_t grid_groups, hipStream_t s);
// *out = within-block covariance of the n_off discretization offsets.
hipError_t launch_block_cbb(int D, bool lean, const DevModel *model, const double *offs,
                            int32_t n_off, double *out, hipStream_t s);
// FP32-refined kriging (lean models, simple / ordinary kriging, no extended
// systems): hipErrorInvalidValue for other shapes.
hipError_t launch_krige_fp32r(int D, int NB, int M, const SolveParams &p, int64_t grid_groups,
                              hipStream_t s);
// FP32-refined Vecchia batch (lean models).
hipError_t launch_vecchia_fp32r(int D, int NB, const VecchiaParams &p, int64_t n_blocks,
                                hipStream_t s);
template <int D, int NB, int M>
hipError_t launch_krige_fp32r_inst(const SolveParams &p, int64_t grid_groups, hipStream_t s);
template <int D, int NB>
hipError_t launch_vecchia_fp32r_inst(const VecchiaParams &p, int64_t n_blocks, hipStream_t s);
// Lanes per system of the FP32-refined kernels.
constexpr int fp32r_group(int NB) { return NB == 16 ? 4 : (NB == 32 ? 16 : 32); }
hipError_t launch_iota_nbr(int32_t *nbr, int32_t *count, int64_t T, int32_t n, int32_t stride,
                           hipStream_t s);
hipError_t launch_zero_i32(int32_t *p, int64_t n, hipStream_t s);
hipError_t launch_fill_f64(double *p, int64_t n, double value, hipStream_t s);
// *dst = min(*src, cap) on the device (stream-ordered)
hipError_t launch_min_i32(const int32_t *src, int32_t cap, int32_t *dst, hipStream_t s);
// perm: processing order of the targets (NULL: index order); n_dev: device
// count of perm's entries (NULL: T); skip: per-target mask of targets to
// leave out (NULL: none).
hipError_t launch_bucket(const int32_t *count, const int32_t *perm, int64_t T, int32_t *lists,
                         int32_t *lens, int64_t list_cap, hipStream_t s,
                         const int32_t *n_dev = nullptr, const uint8_t *skip = nullptr);
// out = the entries t of perm[0..T) with mark[t] == 0, in perm's order;
// *out_len their number.  tmp: select_workspace(T) bytes.
size_t select_workspace(int64_t T);
hipError_t launch_select_unmarked(const int32_t *perm, const uint8_t *mark, int64_t T,
                                  int32_t *out, int32_t *out_len, void *tmp, size_t tmp_bytes,
                                  hipStream_t s);
hipError_t launch_vecchia(int D, int NB, int variant, bool lean, const VecchiaParams &p,
                          int64_t n_blocks, hipStream_t s);
// Vecchia log-likelihood gradient (partial holds 4 columns per model:
// loglik, d/dnugget, d/dpsill, d/drange); NB 16 or 32.
hipError_t launch_vecchia_grad(int D, int NB, const VecchiaParams &p, int64_t n_blocks,
                               hipStream_t s);
hipError_t launch_vecchia_reduce(const double *partial, int64_t n_blocks, int32_t n_models,
                                 double *out, hipStream_t s);

// Per-instantiation launchers (explicit instantiations live in generated
// translation units, one per template argument set, built in parallel).
template <int D, int NB, int M, int G, bool LEAN, bool X>
hipError_t launch_krige_inst(const SolveParams &p, int64_t grid_groups, hipStream_t s);
template <int D, int NB, int G, bool LEAN>
hipError_t launch_vecchia_inst(const VecchiaParams &p, int64_t n_blocks, hipStream_t s);
template <int D, int NB, int G>
hipError_t launch_vecchia_grad_inst(const VecchiaParams &p, int64_t n_blocks, hipStream_t s);

// Models evaluated by one Vecchia launch (host splits larger batches).
constexpr int kMaxModelsPerLaunch = 64;

// Lanes per system (group size G) for a padded system size NB: two
// compiled variants each, picked per configuration by the runtime tuner.
constexpr int kVariants = 2;
constexpr int group_size(int NB, int variant) {
    return NB == 16 ? (variant ? 4 : 2) : NB == 32 ? (variant ? 16 : 8) : (variant ? 64 : 32);
}

}  // namespace gs
