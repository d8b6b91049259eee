// Dispatch of the matrix-core Vecchia kernels (instantiated one per
// generated unit from gs_vecchia_mfma_impl.hpp).
#include "gs_vecchia_mfma.hpp"

namespace gs {

template <int D, bool ANISO, int KIND>
hipError_t launch_vecchia_mfma_inst(const VecchiaParams &p, int64_t n_blocks, hipStream_t s);
template <int D, bool ANISO, int KIND, int M>
hipError_t launch_krige_mfma_inst(const SolveParams &p, int64_t grid_groups, hipStream_t s);
template <int D, int KIND>
hipError_t launch_vecchia_grad_mfma_inst(const VecchiaParams &p, int64_t n_blocks, hipStream_t s);
template <int D, int SIG, int M>
hipError_t launch_krige_mfma64_inst(const SolveParams &p, int64_t grid_groups, hipStream_t s);

int vecchia_mfma_kind(const DevModel *models, int n_models, int max_count, bool *aniso) {
    if (n_models <= 0 || max_count > 32) return -1;
    int kind = -1;
    bool an = false;
    for (int i = 0; i < n_models; i++) {
        const DevModel &m = models[i];
        if (m.ns != 1) return -1;
        int k;
        switch (m.s[0].kind) {
        case K_SPH: k = 0; break;
        case K_MAT15: k = 1; break;
        case K_MAT25: k = 2; break;
        case K_EXP: k = 3; break;
        case K_GAU: k = 4; break;
        default: return -1;
        }
        const bool a = m.s[0].aniso != 0;
        if (i == 0) {
            kind = k;
            an = a;
        } else if (k != kind || a != an) {
            return -1;
        }
    }
    *aniso = an;
    return kind;
}

template <int D, bool AN>
static hipError_t by_kind(int kind, const VecchiaParams &p, int64_t nb, hipStream_t s) {
    switch (kind) {
    case 0: return launch_vecchia_mfma_inst<D, AN, 0>(p, nb, s);
    case 1: return launch_vecchia_mfma_inst<D, AN, 1>(p, nb, s);
    case 2: return launch_vecchia_mfma_inst<D, AN, 2>(p, nb, s);
    case 3: return launch_vecchia_mfma_inst<D, AN, 3>(p, nb, s);
    case 4: return launch_vecchia_mfma_inst<D, AN, 4>(p, nb, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_vecchia_mfma(int D, bool aniso, int kind, const VecchiaParams &p,
                               int64_t n_blocks, hipStream_t s) {
    if (p.points_per_block % 4 != 0) return hipErrorInvalidValue;
    if (D == 2) return aniso ? by_kind<2, true>(kind, p, n_blocks, s) : by_kind<2, false>(kind, p, n_blocks, s);
    if (D == 3) return aniso ? by_kind<3, true>(kind, p, n_blocks, s) : by_kind<3, false>(kind, p, n_blocks, s);
    return hipErrorInvalidValue;
}

template <int D, bool AN, int M>
static hipError_t krige_by_kind(int kind, const SolveParams &p, int64_t gg, hipStream_t s) {
    switch (kind) {
    case 0: return launch_krige_mfma_inst<D, AN, 0, M>(p, gg, s);
    case 1: return launch_krige_mfma_inst<D, AN, 1, M>(p, gg, s);
    case 2: return launch_krige_mfma_inst<D, AN, 2, M>(p, gg, s);
    case 3: return launch_krige_mfma_inst<D, AN, 3, M>(p, gg, s);
    case 4: return launch_krige_mfma_inst<D, AN, 4, M>(p, gg, s);
    default: return hipErrorInvalidValue;
    }
}

template <int D, int M>
static hipError_t krige_by_aniso(bool aniso, int kind, const SolveParams &p, int64_t gg, hipStream_t s) {
    return aniso ? krige_by_kind<D, true, M>(kind, p, gg, s) : krige_by_kind<D, false, M>(kind, p, gg, s);
}

hipError_t launch_krige_mfma(int D, bool aniso, int kind, int M, const SolveParams &p,
                             int64_t grid_groups, hipStream_t s) {
    if (D == 2) return M == 0 ? krige_by_aniso<2, 0>(aniso, kind, p, grid_groups, s)
                              : (M == 1 ? krige_by_aniso<2, 1>(aniso, kind, p, grid_groups, s) : hipErrorInvalidValue);
    if (D == 3) return M == 0 ? krige_by_aniso<3, 0>(aniso, kind, p, grid_groups, s)
                              : (M == 1 ? krige_by_aniso<3, 1>(aniso, kind, p, grid_groups, s) : hipErrorInvalidValue);
    return hipErrorInvalidValue;
}

static int lean_code(int32_t kind) {
    switch (kind) {
    case K_SPH: return 0;
    case K_MAT15: return 1;
    case K_MAT25: return 2;
    case K_EXP: return 3;
    case K_GAU: return 4;
    default: return -1;
    }
}

int krige_mfma64_sig(const DevModel &m, int D) {
    if (m.ns == 1) {
        const int k = lean_code(m.s[0].kind);
        return k < 0 ? -1 : (m.s[0].aniso ? 8 : 0) + k;
    }
    if (m.ns == 2 && m.s[0].kind == K_EXP && m.s[1].kind == K_SPH && m.s[0].aniso == m.s[1].aniso) {
        if (!m.s[0].aniso) return 17;
        for (int i = 0; i < D * D; i++)
            if (m.s[0].m[i] != m.s[1].m[i]) return -1;
        return 16;
    }
    return -1;
}

// The instantiated kernels (one generated unit each): every signature for
// M in {0, 1, 3, 6} in 2-D and {0, 1, 4} in 3-D (simple, ordinary, linear
// and, in 2-D, quadratic drift; the border holds at most six drift rows).
static bool k64_m(int D, int M) {
    return D == 2 ? (M == 0 || M == 1 || M == 3 || M == 6) : (D == 3 && (M == 0 || M == 1 || M == 4));
}

static bool k64_sig(int sig) { return (sig >= 0 && sig <= 4) || (sig >= 8 && sig <= 12) || sig == 16 || sig == 17; }

bool has_krige_mfma64(int D, int sig, int M) { return k64_m(D, M) && k64_sig(sig); }

template <int D, int M>
static hipError_t k64_by_sig(int sig, const SolveParams &p, int64_t gg, hipStream_t s) {
    switch (sig) {
    case 0: return launch_krige_mfma64_inst<D, 0, M>(p, gg, s);
    case 1: return launch_krige_mfma64_inst<D, 1, M>(p, gg, s);
    case 2: return launch_krige_mfma64_inst<D, 2, M>(p, gg, s);
    case 3: return launch_krige_mfma64_inst<D, 3, M>(p, gg, s);
    case 4: return launch_krige_mfma64_inst<D, 4, M>(p, gg, s);
    case 8: return launch_krige_mfma64_inst<D, 8, M>(p, gg, s);
    case 9: return launch_krige_mfma64_inst<D, 9, M>(p, gg, s);
    case 10: return launch_krige_mfma64_inst<D, 10, M>(p, gg, s);
    case 11: return launch_krige_mfma64_inst<D, 11, M>(p, gg, s);
    case 12: return launch_krige_mfma64_inst<D, 12, M>(p, gg, s);
    case 16: return launch_krige_mfma64_inst<D, 16, M>(p, gg, s);
    case 17: return launch_krige_mfma64_inst<D, 17, M>(p, gg, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_krige_mfma64(int D, int sig, int M, const SolveParams &p, int64_t grid_groups,
                               hipStream_t s) {
    if (!has_krige_mfma64(D, sig, M)) return hipErrorInvalidValue;
    if (D == 2) {
        switch (M) {
        case 0: return k64_by_sig<2, 0>(sig, p, grid_groups, s);
        case 1: return k64_by_sig<2, 1>(sig, p, grid_groups, s);
        case 3: return k64_by_sig<2, 3>(sig, p, grid_groups, s);
        default: return k64_by_sig<2, 6>(sig, p, grid_groups, s);
        }
    }
    switch (M) {
    case 0: return k64_by_sig<3, 0>(sig, p, grid_groups, s);
    case 1: return k64_by_sig<3, 1>(sig, p, grid_groups, s);
    default: return k64_by_sig<3, 4>(sig, p, grid_groups, s);
    }
}

template <int D>
static hipError_t grad_by_kind(int kind, const VecchiaParams &p, int64_t nb, hipStream_t s) {
    switch (kind) {
    case 0: return launch_vecchia_grad_mfma_inst<D, 0>(p, nb, s);
    case 1: return launch_vecchia_grad_mfma_inst<D, 1>(p, nb, s);
    case 2: return launch_vecchia_grad_mfma_inst<D, 2>(p, nb, s);
    case 3: return launch_vecchia_grad_mfma_inst<D, 3>(p, nb, s);
    case 4: return launch_vecchia_grad_mfma_inst<D, 4>(p, nb, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_vecchia_grad_mfma(int D, int kind, const VecchiaParams &p, int64_t n_blocks,
                                    hipStream_t s) {
    if (p.points_per_block % 4 != 0) return hipErrorInvalidValue;
    if (D == 2) return grad_by_kind<2>(kind, p, n_blocks, s);
    if (D == 3) return grad_by_kind<3>(kind, p, n_blocks, s);
    return hipErrorInvalidValue;
}

}  // namespace gs
