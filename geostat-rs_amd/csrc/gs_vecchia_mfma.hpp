// The Vecchia likelihood batch and the 17-32-neighbour simple / ordinary
// kriging solve on the matrix cores (gs_vecchia_mfma_impl.hpp), for models
// of one lean structure (the same kind and anisotropy flag across a batch).
#pragma once

#include "gs_internal.hpp"

namespace gs {

// kind code of the kernel for a batch (0 spherical, 1 Matern 3/2, 2 Matern
// 5/2, 3 exponential, 4 Gaussian), or -1 when the batch does not qualify
int vecchia_mfma_kind(const DevModel *models, int n_models, int max_count, bool *aniso);

hipError_t launch_vecchia_mfma(int D, bool aniso, int kind, const VecchiaParams &p,
                               int64_t n_blocks, hipStream_t s);

// Simple (M = 0) or ordinary (M = 1) kriging with 17-32 neighbours and one
// lean structure on the matrix cores (kind as vecchia_mfma_kind);
// grid_groups systems in flight, four per wavefront.
hipError_t launch_krige_mfma(int D, bool aniso, int kind, int M, const SolveParams &p,
                             int64_t grid_groups, hipStream_t s);

// Covariance signature of a model for the 33-64-neighbour matrix-core
// kriging kernel: 0-4 one isotropic lean structure of that kind (as
// vecchia_mfma_kind), 8-12 one anisotropic, 16 / 17 an exponential and a
// spherical structure (in that order) sharing one anisotropic / isotropic
// frame; -1 when the model has none.
int krige_mfma64_sig(const DevModel &m, int D);

// Whether the library holds that kernel for (D, signature, M).
bool has_krige_mfma64(int D, int sig, int M);

// Simple, ordinary or universal kriging with 33-64 neighbours on the
// matrix cores, one system per wavefront, grid_groups wavefronts.
hipError_t launch_krige_mfma64(int D, int sig, int M, const SolveParams &p, int64_t grid_groups,
                               hipStream_t s);

// The log-likelihood gradient batch (vecchia_grad_kernel's quantities) on
// the matrix cores: one isotropic structure of the same kind for every
// model, conditioning sets of at most 32 points, points_per_block % 4 == 0.
hipError_t launch_vecchia_grad_mfma(int D, int kind, const VecchiaParams &p, int64_t n_blocks,
                                    hipStream_t s);

}  // namespace gs
