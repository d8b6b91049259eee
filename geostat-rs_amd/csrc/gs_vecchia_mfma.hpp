// The Vecchia likelihood batch on the matrix cores (gs_vecchia_mfma_impl.hpp),
// for conditioning sets of at most 32 points and batches whose models are
// all one lean structure of the same kind and anisotropy flag.
#pragma once

#include "gs_internal.hpp"

namespace gs {

// kind code of the kernel for a batch (0 spherical, 1 Matern 3/2, 2 Matern
// 5/2, 3 exponential, 4 Gaussian), or -1 when the batch does not qualify
int vecchia_mfma_kind(const DevModel *models, int n_models, int max_count, bool *aniso);

hipError_t launch_vecchia_mfma(int D, bool aniso, int kind, const VecchiaParams &p,
                               int64_t n_blocks, hipStream_t s);

}  // namespace gs
